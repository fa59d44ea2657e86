"""Batched evolutions: one launch per RK4 stage over many uneven grids.

Each member of a batch must end bit-identical to the same grid stepped alone
(and to the oracle), whatever its neighbours in the batch are."""
import numpy as np
import pytest

import oracle
from kg_amd.batch import MemberSpec, batch_step, make_member
from kg_amd.initial_data import random_lm_coeffs

pytestmark = pytest.mark.gpu


def _specs():
    # uneven sizes, masses, black holes; small enough for the oracle
    return [
        MemberSpec(24, 8, 16, 0.1, 1.0, 1.8, 1.8 + 24 * 0.35, 5.0, 1.5, random_lm_coeffs(1, 2, 0.3), 1),
        MemberSpec(40, 16, 32, 0.4, 0.6, 1.08, 1.08 + 40 * 0.3, 6.0, 1.5, random_lm_coeffs(2, 3, 0.3), 2),
        MemberSpec(17, 6, 12, 0.0, 1.9, 3.42, 3.42 + 17 * 0.4, 7.0, 1.2, random_lm_coeffs(3, 2, 0.3), 3),
        MemberSpec(70, 8, 20, 0.25, 1.3, 2.34, 2.34 + 70 * 0.2, 9.0, 2.0, random_lm_coeffs(4, 4, 0.3), 4),
    ]


@pytest.mark.parametrize("nsteps", [1, 4, 5])
def test_batch_equals_each_member_alone(nsteps):
    specs = _specs()
    batch = [make_member(s) for s in specs]
    alone = [make_member(s) for s in specs]
    print("batch of", len(batch), "nsteps", nsteps, flush=True)
    batch_step(batch, nsteps)
    for m in alone:
        m.step(nsteps)
    for q, (b, a) in enumerate(zip(batch, alone)):
        assert np.array_equal(b.u_soa(), a.u_soa()), q
        assert b.t == a.t


def test_batch_matches_oracle_and_interleaves_with_single_steps():
    specs = _specs()
    batch = [make_member(s) for s in specs]
    u0 = [b.u_soa()[0] for b in batch]
    batch[1].step(1)          # members need not share their buffer parity
    batch_step(batch, 3)
    batch[2].step(2)
    batch_step(batch, 2)
    nsteps = [5, 6, 7, 5]
    for q, s in enumerate(specs):
        sim = oracle.Sim(s.Nr, s.Ntheta + 1, s.Nphi + 1, 2, s.xmin, s.xmax, "Kerr_Schild", 0.0, s.M_bh, s.m,
                         "outflow", 0.25)
        sim.u[...] = u0[q]
        sim.step("rk4", nsteps[q])
        assert np.array_equal(batch[q].u_soa()[0], sim.u), q


def test_fd4_batch_equals_each_member_alone_and_the_oracle():
    """The 4th-order operator set in batched launches: every member equals its
    solo run and the oracle, bit for bit; a batch may not mix operator sets."""
    specs = _specs()
    batch = [make_member(s, operators="fd4") for s in specs]
    alone = [make_member(s, operators="fd4") for s in specs]
    u0 = [b.u_soa()[0] for b in batch]
    print("fd4 batch", flush=True)
    batch_step(batch, 3)
    batch[1].step(1)
    batch_step(batch, 2)
    for a in alone:
        a.step(5)
    alone[1].step(1)
    for q, s in enumerate(specs):
        assert np.array_equal(batch[q].u_soa(), alone[q].u_soa()), q
        sim = oracle.Sim(s.Nr, s.Ntheta + 1, s.Nphi + 1, 2, s.xmin, s.xmax, "Kerr_Schild", 0.0, s.M_bh, s.m,
                         "outflow", 0.25)
        sim.set_order(4)
        sim.u[...] = u0[q]
        sim.step("rk4", 6 if q == 1 else 5)
        assert np.array_equal(batch[q].u_soa()[0], sim.u), q
    from kg_amd import KGError
    mixed = [make_member(specs[0], operators="fd4"), make_member(specs[1])]
    with pytest.raises(KGError):
        batch_step(mixed, 1)


@pytest.mark.parametrize("scheme,operators", [("rk3", "ref2"), ("pirk2", "ref2"), ("pirk3", "ref2"), ("rk3", "fd4"),
                                              ("rk4", "ref2")])
def test_every_scheme_batches_like_each_member_alone(scheme, operators):
    """Batched launches for every integrator: members of uneven sizes (one a
    complex state), a member stepped alone in between so the stale-Pi and
    buffer parities differ across the batch, odd and even step counts; every
    member ends bit-identical to its solo run."""
    specs = _specs()
    cx = [False, True, False, False]
    batch = [make_member(s, scheme=scheme, operators=operators, is_complex=c) for s, c in zip(specs, cx)]
    alone = [make_member(s, scheme=scheme, operators=operators, is_complex=c) for s, c in zip(specs, cx)]
    print("batch", scheme, operators, flush=True)
    batch_step(batch, 3)
    batch[2].step(1)
    batch_step(batch, 2)
    for q, a in enumerate(alone):
        a.step(6 if q == 2 else 5)
    for q in range(len(specs)):
        assert np.array_equal(batch[q].u_soa(), alone[q].u_soa()), (scheme, q)
        assert batch[q].t == alone[q].t
        if scheme in ("pirk2", "pirk3"):  # the stale-Pi stage arrays carry over too
            assert np.array_equal(batch[q].u_stage_soa(0), alone[q].u_stage_soa(0)), (scheme, q)
        batch[q].close()
        alone[q].close()


@pytest.mark.parametrize("scheme,operators,wg", [("rk4", "fd4", 1), ("rk4", "ref2", 2), ("pirk2", "ref2", 1),
                                                 ("rk3", "fd4", 3)])
def test_persistent_batch_launches_equal_each_member_alone(scheme, operators, wg):
    """Persistent batched launches (wg workgroups per CU march all the tiles
    in turn; with two members of the population's smallest class, 256 x 16 x
    32, there are more tiles than workgroups): every member bit-identical to
    its solo run, and switching back to one workgroup per tile continues
    identically."""
    dr = 58.2 / 256
    specs = _specs() + [
        MemberSpec(256, 16, 32, 0.3, 0.8, 1.44, 1.44 + 256 * dr, 20.0, 1.5, random_lm_coeffs(5, 4, 1e-3), 5),
        MemberSpec(256, 16, 32, 0.1, 1.6, 2.88, 2.88 + 256 * dr, 30.0, 1.5, random_lm_coeffs(6, 4, 1e-3), 6),
    ]
    batch = [make_member(s, scheme=scheme, operators=operators) for s in specs]
    alone = [make_member(s, scheme=scheme, operators=operators) for s in specs]
    print("persistent batch", scheme, operators, wg, flush=True)
    batch[0].set_batch_persistent(wg)
    batch_step(batch, 3)
    batch[0].set_batch_persistent(0)
    batch_step(batch, 2)
    for q, a in enumerate(alone):
        a.step(5)
        assert np.array_equal(batch[q].u_soa(), a.u_soa()), (scheme, q)
        a.close()
        batch[q].close()


def test_pirk2_batch_matches_the_oracle():
    """The reference's live integrator in batched launches against the oracle."""
    specs = _specs()
    batch = [make_member(s, scheme="pirk2") for s in specs]
    u0 = [b.u_soa()[0] for b in batch]
    print("pirk2 batch vs oracle", flush=True)
    batch_step(batch, 4)
    for q, s in enumerate(specs):
        sim = oracle.Sim(s.Nr, s.Ntheta + 1, s.Nphi + 1, 2, s.xmin, s.xmax, "Kerr_Schild", 0.0, s.M_bh, s.m,
                         "outflow", 0.25)
        sim.u[...] = u0[q]
        sim.step("pirk2", 4)
        assert np.array_equal(batch[q].u_soa()[0], sim.u), q


def test_batch_refuses_mixed_schemes():
    from kg_amd import KGError
    from kg_amd import _capi as capi
    specs = _specs()
    mixed = [make_member(specs[0], scheme="rk4"), make_member(specs[1], scheme="pirk2")]
    with pytest.raises(KGError) as ei:
        batch_step(mixed, 1)
    assert ei.value.status == capi.KG_ERR_UNSUPPORTED


def test_batch_autotuner_picks_a_candidate_and_results_stay_solo_identical():
    """autotune_batch times its candidates on the batch itself (advancing it
    1 + steps steps per candidate) and leaves every member at the pick; the
    members still match their solo runs step for step afterwards."""
    from kg_amd.autotune import autotune_batch
    specs = _specs()
    batch = [make_member(s, operators="fd4") for s in specs]
    alone = [make_member(s, operators="fd4") for s in specs]
    print("batch autotune", flush=True)
    kc, timings = autotune_batch(batch, candidates=(2, 8, 64), steps=1)
    assert kc in (2, 8, 64) and set(timings) == {2, 8, 64}
    assert all(t > 0 for t in timings.values())
    batch_step(batch, 2)
    for a in alone:
        a.step(3 * 2 + 2)  # three candidates x (1 warm-up + 1 timed) + 2
    for q in range(len(specs)):
        assert np.array_equal(batch[q].u_soa(), alone[q].u_soa()), q
        batch[q].close()
        alone[q].close()
