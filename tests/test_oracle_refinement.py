"""Two-level refinement on the oracle (oracle/refinement.py), pinned by
mathematics.  The reference has no refinement levels, so nothing of its own
pins this; the checks are:

- the buffer refill is exact on what it must reproduce: quintic polynomials
  in r (6-point interpolation at midpoints, copies at coincident points) and
  quadratics in time (the half-step weights -1/8, 3/4, 3/8; linears for the
  two-time 1/2, 1/2 start);
- restriction injects the fine values at coincident radii;
- a refined run converges at the operator set's order on both levels when
  the base resolution is doubled twice (dt held fixed, so the integrators'
  time errors are the same at every resolution and cancel in the differences:
  ratio ~4 for the 2nd-order set with PIRK2, PIRK3 and RK3, ~16 for fd4 with
  RK3 and RK4);
- inside the patch a refined run is as accurate as a uniform run at the fine
  resolution, and far more accurate than one at the base resolution.
"""
import math

import numpy as np
import pytest

from kg_amd.initial_data import shell_state
from oracle import Sim
from oracle.refinement import RefinedSim

XMIN, XMAX, RP = 1.8, 41.8, 10.0


def _refined(Nr, order, dt0=0.025, Ny=5, Nz=9, coeffs=None, r0=8.0, w=1.5):
    R = RefinedSim(Nr, Ny, Nz, 2, XMIN, XMAX, "Kerr_Schild", 0.0, 1.0, 0.2, "outflow", 0.25, RP, dt=dt0, order=order)
    for lv in (R.base, R.fine):
        u0 = shell_state(lv.Nx, lv.Ny, lv.Nz, 2, XMIN, XMIN + lv.Nx * lv.dx, r0, w, 1.0, coeffs or {(0, 0): 1.0})
        lv.u[...] = lv.fill(u0.copy()) if order == 4 else u0
    return R


def test_buffer_refill_is_exact_for_quintics_in_r_and_quadratics_in_t():
    R = _refined(64, 2)
    xb = R.base.table("x")[R.c_lo:R.c_lo + R.ncol]
    xf = R.fine.table("x")
    P = lambda x: 0.3 + 0.1 * x - 0.02 * x ** 2 + 3e-3 * x ** 3 - 2e-4 * x ** 4 + 5e-6 * x ** 5
    Q = lambda t: 1.0 + 0.4 * t - 0.3 * t ** 2
    shape = R.base.u[..., :R.ncol].shape
    slots = [np.broadcast_to(P(xb) * Q(t), shape).copy() for t in (-1.0, 0.0, 1.0)]
    ni = R.fine.ni
    buf = slice(ni - R.B, ni)
    scale = np.abs(P(xf[buf])).max()
    # copy of the base time t
    R._prolong([slots[1]], [1.0])
    assert np.abs(R.fine.u[..., buf] - P(xf[buf]) * Q(0.0)).max() < 1e-13 * scale
    # the half step through three base times
    R._prolong(slots, [-0.125, 0.75, 0.375])
    assert np.abs(R.fine.u[..., buf] - P(xf[buf]) * Q(0.5)).max() < 1e-13 * scale
    # with two known times: exact for linears in t
    L = lambda t: 1.0 + 0.4 * t
    lin = [np.broadcast_to(P(xb) * L(t), shape).copy() for t in (0.0, 1.0)]
    R._prolong(lin, [0.5, 0.5])
    assert np.abs(R.fine.u[..., buf] - P(xf[buf]) * L(0.5)).max() < 1e-13 * scale
    # and a sextic in r is not reproduced: the rule is exactly 6-point
    S = lambda x: (x - 5.0) ** 6 * 1e-6
    R._prolong([np.broadcast_to(S(xb), shape).copy()], [1.0])
    assert np.abs(R.fine.u[..., buf] - S(xf[buf])).max() > 1e-12


def test_restriction_injects_fine_values():
    R = _refined(64, 2)
    R.step("rk4", 3)
    nc = R.I - 2
    assert np.array_equal(R.base.u[..., :nc], R.fine.u[..., 0:2 * nc:2])
    # the fine level's coincident points sit at the base radii
    xb, xf = R.base.table("x"), R.fine.table("x")
    assert np.abs(xb[:nc] - xf[0:2 * nc:2]).max() < 1e-12
    # and the patch edge is the base point nearest r_patch
    assert abs(xb[R.I] - RP) <= 0.5 * (xb[1] - xb[0])


@pytest.mark.parametrize("scheme,order", [("pirk2", 2), ("pirk3", 2), ("rk3", 2), ("rk3", 4), ("rk4", 4)])
def test_refined_run_converges_at_the_operator_sets_order(scheme, order):
    """An l = 0 pulse (r0 = 8, width 1.5, inside the patch) crossing r_patch = 10
    outward over T = 6; base Nr = 160, 320, 640 on [1.8, 41.8], dt = 0.025.
    Compared on r in [3, 30] (base) and [3, 10] (fine), away from the
    excision boundary's extrapolation."""
    lo, hi = (3.4, 5.2) if order == 2 else (13.0, 21.0)
    runs = []
    for N in (160, 320, 640):
        R = _refined(N, order)
        R.step(scheme, 240)
        runs.append((R.base.u.copy(), R.fine.u.copy(), R.I))
    dx0 = (XMAX - XMIN) / 160
    ib = np.arange(int(1.2 / dx0) + 1, int(28.2 / dx0))
    jf = np.arange(int(1.2 / (dx0 / 2)) + 1, 2 * runs[0][2] + 1)
    for level, idx in ((0, ib), (1, jf)):
        a, b, c = runs[0][level][..., idx], runs[1][level][..., 2 * idx], runs[2][level][..., 4 * idx]
        for f in (0, 1, 4):
            e1, e2 = np.abs(a[f] - b[f]).max(), np.abs(b[f] - c[f]).max()
            assert lo < e1 / e2 < hi, (scheme, order, "level", level, "field", f, e1, e2)


def test_patch_is_as_accurate_as_the_fine_resolution():
    """RK4 with fd4, a pulse well inside the patch (r0 = 6, width 1: its tail
    at r_patch is below 1e-6) over T = 2: inside the patch a refined run at
    base resolution N is within 10 % of a uniform run at 2N (errors against a
    uniform run at 8N) and at least 10 times more accurate than a uniform run
    at N."""
    T, dt0 = 2.0, 0.025

    def uniform(N, dt):
        s = Sim(N, 5, 9, 2, XMIN, XMAX, "Kerr_Schild", 0.0, 1.0, 0.2, "outflow", 0.25, dt)
        s.set_order(4)
        s.u[...] = s.fill(shell_state(N, 5, 9, 2, XMIN, XMAX, 6.0, 1.0, 1.0, {(0, 0): 1.0}))
        s.step("rk4", int(round(T / dt)))
        return s.u.copy()

    R = _refined(160, 4, dt0, r0=6.0, w=1.0)
    R.step("rk4", int(round(T / dt0)))
    ref = uniform(1280, dt0 / 2)
    u1, u2 = uniform(160, dt0), uniform(320, dt0 / 2)
    dxf = (XMAX - XMIN) / 320
    jf = np.arange(int(1.2 / dxf) + 1, 2 * R.I + 1, 2)  # even fine columns on r in [3, r_patch]: base points too
    ref_f = ref[..., 4 * jf]
    e_ref = np.abs(R.fine.u[..., jf] - ref_f).max()
    e_2N = np.abs(u2[..., jf] - ref_f).max()
    e_N = np.abs(u1[..., jf // 2] - ref_f).max()
    assert e_ref < 1.1 * e_2N, (e_ref, e_2N)
    assert e_N > 10.0 * e_ref, (e_N, e_ref)
    assert math.isfinite(e_ref)
