"""The Sommerfeld outer boundary on the GPU against the oracle.

The outer column is evolved with the radiative condition on Pi (the oracle's
statement, pinned in test_oracle_sommerfeld.py by the exact outgoing wave);
the GPU must reproduce the oracle bit for bit for every scheme and operator
set, real and complex states, alone and batched, and on the small workload
exactly as specified (Sommerfeld outer boundary, 2000 RK4 steps).
"""
import numpy as np
import pytest

import oracle
from kg_amd import KGError, LinesMethod
from kg_amd.batch import MemberSpec, batch_step
from kg_amd.initial_data import random_lm_coeffs, shell_state

pytestmark = pytest.mark.gpu

CASES = {
    "ks_spin": dict(Nx=20, Ny=12, Nz=14, xmin=1.9, xmax=12.0, metric="Kerr_Schild", a=0.6, M_bh=1.0, m=0.3),
    "schw_ef": dict(Nx=18, Ny=10, Nz=21, xmin=2.5, xmax=9.0, metric="Schw_EF", a=0.0, M_bh=1.0, m=0.1),
    "flat": dict(Nx=24, Ny=9, Nz=17, xmin=1.0, xmax=6.0, metric="Kerr_Schild", a=0.0, M_bh=0.0, m=0.0),
}


def _data(p, seed):
    rng = np.random.default_rng(seed)
    u0 = shell_state(p["Nx"], p["Ny"], p["Nz"], 2, p["xmin"], p["xmax"], 0.6 * p["xmax"], 1.5, 1.0,
                     random_lm_coeffs(seed, lmax=3, std=0.3))
    u0[4] = 0.1 * rng.standard_normal(u0[4].shape)
    u0 += 1e-3 * rng.standard_normal(u0.shape)
    return u0


def _sim(p, order):
    sim = oracle.Sim(p["Nx"], p["Ny"], p["Nz"], 2, p["xmin"], p["xmax"], p["metric"], p["a"], p["M_bh"], p["m"],
                     "sommerfeld", 0.25)
    sim.set_order(order)
    return sim


def _gpu(p, scheme, operators, cplx=False):
    return LinesMethod(Nx=p["Nx"], Ny=p["Ny"], Nz=p["Nz"], xmin=p["xmin"], xmax=p["xmax"], CFL=0.25,
                       boundary_type_rmax="sommerfeld", metric=p["metric"], a=p["a"], M_bh=p["M_bh"], m=p["m"],
                       scheme=scheme, operators=operators, is_complex=cplx)


SCHEMES = [("pirk2", "ref2"), ("pirk3", "ref2"), ("rk3", "ref2"), ("rk4", "ref2"), ("rk3", "fd4"), ("rk4", "fd4")]


@pytest.mark.parametrize("scheme,operators", SCHEMES)
@pytest.mark.parametrize("name", list(CASES))
def test_sommerfeld_steps_match_oracle(scheme, operators, name):
    p = CASES[name]
    print("sommerfeld", name, scheme, operators, flush=True)
    sim = _sim(p, 4 if operators == "fd4" else 2)
    u0 = _data(p, 5)
    if operators == "fd4":
        u0 = sim.fill(u0.copy())
    sim.u[...] = u0
    gpu = _gpu(p, scheme, operators)
    gpu.set_u_soa(u0[None])
    sim.step(scheme, 7)
    gpu.step(7)
    got = gpu.u_soa()[0]
    assert np.isfinite(got).all()
    assert np.array_equal(got, sim.u), (name, scheme, operators)
    if scheme in ("pirk2", "pirk3"):
        assert np.array_equal(gpu.u_stage_soa(0)[0], sim.u_stage)
        if scheme == "pirk3":
            assert np.array_equal(gpu.u_stage_soa(1)[0], sim.u_stage_2)
    gpu.close()


@pytest.mark.parametrize("scheme,operators", [("pirk2", "ref2"), ("rk4", "fd4")])
def test_sommerfeld_complex_state(scheme, operators):
    p = CASES["ks_spin"]
    order = 4 if operators == "fd4" else 2
    parts = [_data(p, 6), _data(p, 7)]
    if operators == "fd4":  # ghost-filled data (the fill works in place and is not idempotent: once)
        s = _sim(p, order)
        parts = [s.fill(x.copy()) for x in parts]
        s.close()
    want = []
    for u0 in parts:
        sim = _sim(p, order)
        sim.u[...] = u0
        sim.step(scheme, 5)
        want.append(sim.u.copy())
        sim.close()
    gpu = _gpu(p, scheme, operators, cplx=True)
    gpu.set_u_soa(np.stack(parts))
    gpu.step(5)
    got = gpu.u_soa()
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    gpu.close()


def _member(spec, operators):
    lm = LinesMethod(Nx=spec.Nr, Ny=spec.Ntheta + 1, Nz=spec.Nphi + 1, xmin=spec.xmin, xmax=spec.xmax, CFL=0.25,
                     m=spec.m, M_bh=spec.M_bh, scheme="rk4", operators=operators, boundary_type_rmax="sommerfeld")
    u0 = shell_state(spec.Nr, spec.Ntheta + 1, spec.Nphi + 1, 2, spec.xmin, spec.xmax, spec.r0, spec.width, 1.0,
                     spec.coeffs)
    lm.set_u_soa(u0[None])
    lm.boundary()
    return lm


@pytest.mark.parametrize("operators", ["ref2", "fd4"])
def test_sommerfeld_batch_equals_alone_and_oracle(operators):
    specs = [
        MemberSpec(24, 8, 16, 0.1, 1.0, 1.8, 1.8 + 24 * 0.35, 5.0, 1.5, random_lm_coeffs(1, 2, 0.3), 1),
        MemberSpec(40, 16, 32, 0.4, 0.6, 1.08, 1.08 + 40 * 0.3, 6.0, 1.5, random_lm_coeffs(2, 3, 0.3), 2),
        MemberSpec(17, 6, 12, 0.0, 1.9, 3.42, 3.42 + 17 * 0.4, 7.0, 1.2, random_lm_coeffs(3, 2, 0.3), 3),
    ]
    batch = [_member(s, operators) for s in specs]
    alone = [_member(s, operators) for s in specs]
    u0 = [b.u_soa()[0] for b in batch]
    batch_step(batch, 3)
    batch_step(batch, 2, noalloc=True)
    for a in alone:
        a.step(5)
    for q, s in enumerate(specs):
        assert np.array_equal(batch[q].u_soa(), alone[q].u_soa()), q
        sim = oracle.Sim(s.Nr, s.Ntheta + 1, s.Nphi + 1, 2, s.xmin, s.xmax, "Kerr_Schild", 0.0, s.M_bh, s.m,
                         "sommerfeld", 0.25)
        sim.set_order(4 if operators == "fd4" else 2)
        sim.u[...] = u0[q]
        sim.step("rk4", 5)
        assert np.array_equal(batch[q].u_soa()[0], sim.u), q
    # a batch may not mix Sommerfeld with the filled outer boundaries
    other = LinesMethod(Nx=24, Ny=9, Nz=17, xmin=1.8, xmax=10.2, scheme="rk4", operators=operators)
    with pytest.raises(KGError):
        batch_step([batch[0], other], 1)
    for m in batch + alone + [other]:
        m.close()


@pytest.mark.parametrize("operators", ["ref2", "fd4"])
def test_outgoing_flat_space_pulse_matches_oracle(operators):
    """The oracle's exact-solution run (outgoing f(t-r)/r through the outer
    boundary) on the GPU: bit-identical after the pulse has left."""
    Nr = 160
    sim = oracle.Sim(Nr, 5, 9, 2, 5.0, 25.0, "Kerr_Schild", 0.0, 0.0, 0.0, "sommerfeld", 0.25)
    sim.set_order(4 if operators == "fd4" else 2)
    x = sim.table("x")
    u = 0.0 - x + 15.0
    f = np.exp(-(u ** 2) / 4.0)
    fp = -2.0 * u / 4.0 * f
    u0 = np.zeros(sim.shape)
    u0[0], u0[1], u0[4] = f / x, -fp / x - f / x ** 2, fp / x
    sim.u[...] = u0
    gpu = LinesMethod(Nx=Nr, Ny=5, Nz=9, xmin=5.0, xmax=25.0, CFL=0.25, M_bh=0.0, boundary_type_rmax="sommerfeld",
                      scheme="rk4", operators=operators)
    gpu.set_u_soa(u0[None])
    n = int(round(16.0 / sim.dt))
    sim.step("rk4", n)
    gpu.step(n)
    got = gpu.u_soa()[0]
    assert np.array_equal(got, sim.u)
    assert np.abs(got[0]).max() < 1e-3  # the pulse has left
    gpu.close()
