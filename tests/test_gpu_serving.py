"""The serving loop on the GPU: requests admitted under an HBM budget and
advanced by continuous batching give the same final fields and spectra as
each request run alone through the plain path (bit-identical)."""
import numpy as np
import pytest

from kg_amd.batch import MemberSpec
from kg_amd.initial_data import random_lm_coeffs
from kg_amd.serving import GpuEngine, Request, ServingLoop

pytestmark = pytest.mark.gpu


def _requests():
    specs = [
        MemberSpec(24, 8, 16, 0.1, 1.0, 1.8, 1.8 + 24 * 0.35, 5.0, 1.5, random_lm_coeffs(1, 2, 0.3), 1),
        MemberSpec(40, 16, 32, 0.4, 0.6, 1.08, 1.08 + 40 * 0.3, 6.0, 1.5, random_lm_coeffs(2, 3, 0.3), 2),
        MemberSpec(17, 6, 12, 0.0, 1.9, 3.42, 3.42 + 17 * 0.4, 7.0, 1.2, random_lm_coeffs(3, 2, 0.3), 3),
        MemberSpec(70, 8, 20, 0.25, 1.3, 2.34, 2.34 + 70 * 0.2, 9.0, 2.0, random_lm_coeffs(4, 4, 0.3), 4),
        MemberSpec(33, 16, 32, 0.3, 0.8, 1.44, 1.44 + 33 * 0.25, 5.5, 1.5, random_lm_coeffs(5, 4, 0.3), 5),
        MemberSpec(50, 8, 16, 0.05, 1.6, 2.88, 2.88 + 50 * 0.3, 8.0, 1.5, random_lm_coeffs(6, 1, 0.3), 6),
    ]
    steps = [9, 4, 13, 6, 3, 11]
    arrivals = [0.0, 0.0, 0.002, 0.004, 0.006, 0.05]
    return [Request(i, s, n, a) for i, (s, n, a) in enumerate(zip(specs, steps, arrivals))]


def test_served_results_equal_running_alone():
    eng = GpuEngine(lmax=4, nk=16, kmax=4.0, on_device=False)
    reqs = _requests()
    sizes = sorted(eng.bytes_for(r.spec) for r in reqs)
    budget = sizes[-1] + sizes[-2]  # at most two or three requests resident at once
    loop = ServingLoop(eng, budget, chunk=4)
    print("serving %d requests, budget %d bytes" % (len(reqs), budget), flush=True)
    results, stats = loop.run(reqs)
    assert stats.max_bytes <= budget
    assert stats.joined_running >= 1 and max(stats.batch_sizes) >= 2
    for req, res in zip(reqs, results):
        assert not res.error
        fields, spectrum = eng.run_alone(req)
        assert np.array_equal(res.fields, fields), req.rid
        assert np.array_equal(res.spectrum, spectrum), req.rid
        assert np.isfinite(spectrum).all() and np.abs(spectrum).max() > 0


def test_fd4_served_with_per_chunk_phi_chunks_equal_running_alone():
    """fd4 stencils, the phi chunk chosen per batched call from the active set
    (a tiny workgroup target, so the choice changes as requests come and go):
    still bit-identical to each request run alone."""
    eng = GpuEngine(lmax=4, nk=16, kmax=4.0, on_device=False, operators="fd4", tuning=(8, 8), batch_kc="fill",
                    fill_workgroups=4)
    reqs = _requests()
    sizes = sorted(eng.bytes_for(r.spec) for r in reqs)
    loop = ServingLoop(eng, sizes[-1] + sizes[-2] + sizes[-3], chunk=4)
    print("serving %d fd4 requests, phi chunk per call" % len(reqs), flush=True)
    results, stats = loop.run(reqs)
    assert len(eng.kc_chunks) >= 2, eng.kc_chunks
    for req, res in zip(reqs, results):
        assert not res.error
        fields, spectrum = eng.run_alone(req)
        assert np.array_equal(res.fields, fields), req.rid
        assert np.array_equal(res.spectrum, spectrum), req.rid


def test_pirk2_served_results_equal_running_alone():
    """The reference's live integrator through the serving loop."""
    eng = GpuEngine(lmax=4, nk=16, kmax=4.0, on_device=False, scheme="pirk2")
    reqs = _requests()
    sizes = sorted(eng.bytes_for(r.spec) for r in reqs)
    loop = ServingLoop(eng, sizes[-1] + sizes[-2], chunk=4)
    print("serving %d pirk2 requests" % len(reqs), flush=True)
    results, stats = loop.run(reqs)
    assert stats.joined_running >= 1 and max(stats.batch_sizes) >= 2
    for req, res in zip(reqs, results):
        assert not res.error
        fields, spectrum = eng.run_alone(req)
        assert np.array_equal(res.fields, fields), req.rid
        assert np.array_equal(res.spectrum, spectrum), req.rid
