"""GPU parity and convergence of the fd4 operator set: the reference's
4th-order stencils (tools/finite_differences.f90: x_4 :124-151, y_4 :241-268,
z_4 :340-367, advec_x_4 :922-957) in the right-hand side of
lines_method_lib.f90:671-753, with RK3 and RK4.

The bar is bit equality with the oracle run with the same operator set
(whose stencils are pinned bit-exact to the reference's compiled functions);
the per-field relative error is also bounded so a failure reports its size.
"""
import numpy as np
import pytest

import oracle
from kg_amd import KGError, LinesMethod
from kg_amd import _capi as capi
from kg_amd.initial_data import random_lm_coeffs, shell_state

pytestmark = pytest.mark.gpu

CASES = [
    ("ks_outflow", dict(Nx=24, Ny=9, Nz=17, xmin=1.8, xmax=10.0, metric="Kerr_Schild", a=0.0, M_bh=1.0, m=0.0,
                        bc="outflow")),
    ("ks_spin_inflow_oddphi", dict(Nx=20, Ny=12, Nz=14, xmin=1.9, xmax=12.0, metric="Kerr_Schild", a=0.6, M_bh=1.0,
                                   m=0.3, bc="inflow")),
    ("schw_ef", dict(Nx=18, Ny=10, Nz=21, xmin=2.5, xmax=9.0, metric="Schw_EF", a=0.0, M_bh=1.0, m=0.1,
                     bc="outflow")),
    ("flat_beta_zero", dict(Nx=16, Ny=8, Nz=11, xmin=1.0, xmax=5.0, metric="Kerr_Schild", a=0.0, M_bh=0.0, m=0.2,
                            bc="outflow")),
    ("beta_negative", dict(Nx=17, Ny=7, Nz=10, xmin=1.8, xmax=6.0, metric="Kerr_Schild", a=0.0, M_bh=-0.5, m=0.0,
                           bc="inflow")),
    ("wide_r_tiles", dict(Nx=140, Ny=11, Nz=9, xmin=1.8, xmax=30.0, metric="Kerr_Schild", a=0.3, M_bh=1.0, m=0.2,
                          bc="outflow")),
    ("tall_theta", dict(Nx=70, Ny=37, Nz=12, xmin=1.8, xmax=15.0, metric="Kerr_Schild", a=0.5, M_bh=1.0, m=0.2,
                        bc="sommerfeld")),
    # boxes of 64k + 1, 2, 3 r columns: the last columns run as 4-wide strips
    # of 128 theta rows (two strips on the 139-row box)
    ("r_strip_1col", dict(Nx=126, Ny=9, Nz=9, xmin=1.8, xmax=25.0, metric="Kerr_Schild", a=0.4, M_bh=1.0, m=0.2,
                          bc="outflow")),
    ("r_strip_2col", dict(Nx=127, Ny=8, Nz=10, xmin=1.8, xmax=25.0, metric="Kerr_Schild", a=0.0, M_bh=1.0, m=0.1,
                          bc="inflow")),
    ("r_strip_3col", dict(Nx=128, Ny=10, Nz=9, xmin=2.5, xmax=25.0, metric="Schw_EF", a=0.0, M_bh=1.0, m=0.2,
                          bc="sommerfeld")),
    ("r_strips_tall", dict(Nx=64, Ny=140, Nz=7, xmin=1.8, xmax=12.0, metric="Kerr_Schild", a=0.5, M_bh=1.0, m=0.3,
                           bc="sommerfeld")),
]


def _rel_err(a, b):
    out = []
    for c in range(5):
        den = np.abs(b[c]).max()
        out.append(float(np.abs(a[c] - b[c]).max() / den) if den > 0 else float(np.abs(a[c]).max()))
    return out


def _pair(p, scheme, seed=0, is_complex=False):
    sim = oracle.Sim(p["Nx"], p["Ny"], p["Nz"], 2, p["xmin"], p["xmax"], p["metric"], p["a"], p["M_bh"], p["m"],
                     p["bc"], 0.25)
    sim.set_order(4)
    rng = np.random.default_rng(seed)
    u0 = shell_state(p["Nx"], p["Ny"], p["Nz"], 2, p["xmin"], p["xmax"], 0.5 * (p["xmin"] + p["xmax"]), 1.5, 1.0,
                     random_lm_coeffs(seed, lmax=3, std=0.3))
    u0[4] = 0.1 * rng.standard_normal(u0[4].shape)
    u0 += 1e-3 * rng.standard_normal(u0.shape)
    gpu = LinesMethod(Nx=p["Nx"], Ny=p["Ny"], Nz=p["Nz"], xmin=p["xmin"], xmax=p["xmax"], CFL=0.25,
                      boundary_type_rmax=p["bc"], metric=p["metric"], a=p["a"], M_bh=p["M_bh"], m=p["m"],
                      scheme=scheme, operators="fd4", is_complex=is_complex)
    assert gpu.dt == sim.dt
    return sim, gpu, u0


@pytest.mark.parametrize("scheme", ["rk3", "rk4"])
@pytest.mark.parametrize("name,p", CASES, ids=[c[0] for c in CASES])
def test_fd4_steps_match_oracle(scheme, name, p):
    print("fd4 case", name, scheme, flush=True)
    sim, gpu, u0 = _pair(p, scheme)
    sim.u[...] = u0
    gpu.set_u_soa(u0[None])
    nsteps = 7
    sim.step(scheme, nsteps)
    gpu.step(nsteps)
    got = gpu.u_soa()[0]
    err = _rel_err(got, sim.u)
    assert max(err) <= 1e-13, (name, scheme, err)
    assert np.array_equal(got, sim.u), (name, scheme, "not bit-identical", err)
    gpu.close()


@pytest.mark.parametrize("kc,tj", [(1, 4), (2, 8), (3, 8), (64, 8)])
@pytest.mark.parametrize("case", [1, 6, 9])
def test_fd4_launch_shape_does_not_change_results(kc, tj, case):
    """The fd4 stages run tiles of 8 theta rows whatever tj; the phi chunk
    varies (1, 2, 3 planes and a whole column)."""
    p = CASES[case][1]
    sim, gpu, u0 = _pair(p, "rk4", seed=5)
    sim.u[...] = u0
    gpu.set_u_soa(u0[None])
    gpu.set_tuning(kc, tj)
    sim.step("rk4", 4)
    gpu.step(4)
    assert np.array_equal(gpu.u_soa()[0], sim.u), (kc, tj)
    gpu.close()


@pytest.mark.parametrize("case", [0, 10])
def test_fd4_complex_state(case):
    p = CASES[case][1]
    sim, gpu, ure = _pair(p, "rk4", seed=3, is_complex=True)
    uim = shell_state(p["Nx"], p["Ny"], p["Nz"], 2, p["xmin"], p["xmax"], 6.0, 1.0, 0.3, {(1, 1): 1.0})
    gpu.set_u_soa(np.stack([ure, uim]))
    gpu.step(6)
    got = gpu.u_soa()
    for part, arr in ((0, ure), (1, uim)):
        s2 = oracle.Sim(p["Nx"], p["Ny"], p["Nz"], 2, p["xmin"], p["xmax"], p["metric"], p["a"], p["M_bh"],
                        p["m"], p["bc"], 0.25)
        s2.set_order(4)
        s2.u[...] = arr
        s2.step("rk4", 6)
        assert np.array_equal(got[part], s2.u), part
    gpu.close()


def test_fd4_one_thousand_steps():
    """1000 RK4 steps with the fd4 set on a small grid: still bit-identical."""
    p = dict(Nx=20, Ny=7, Nz=9, xmin=1.8, xmax=12.0, metric="Kerr_Schild", a=0.0, M_bh=1.0, m=0.2, bc="outflow")
    sim, gpu, u0 = _pair(p, "rk4", seed=11)
    u0 = sim.fill(u0)
    sim.u[...] = u0
    gpu.set_u_soa(u0[None])
    print("fd4 1000 steps", flush=True)
    for _ in range(4):
        sim.step("rk4", 250)
        gpu.step(250)
    got = gpu.u_soa()[0]
    assert np.isfinite(got).all() and np.abs(got).max() > 0
    assert max(_rel_err(got, sim.u)) <= 1e-12
    assert np.array_equal(got, sim.u)
    gpu.close()


def _run_pulse(Nx, T, xmin=1.8, xmax=41.8):
    lm = LinesMethod(Nx=Nx, Ny=5, Nz=9, xmin=xmin, xmax=xmax, CFL=0.25, m=0.2, scheme="rk4", operators="fd4")
    u0 = shell_state(Nx, 5, 9, 2, xmin, xmax, 15.0, 2.0, 1.0, {(0, 0): 1.0})
    lm.set_u_soa(u0[None])
    lm.boundary()
    n = int(round(T / lm.dt))
    assert abs(n * lm.dt - T) < 1e-9
    lm.step(n)
    u = lm.u_soa()[0]
    lm.close()
    return u


def test_fd4_rk4_converges_at_fourth_order():
    """Self-convergence at three resolutions (dt proportional to dr; a
    spherically symmetric massive pulse on Schwarzschild-KS): the ratio of
    successive differences at the common radii approaches 2^4 = 16."""
    T = 6.0
    u = [_run_pulse(N, T) for N in (160, 320, 640)]
    dx = 40.0 / 160
    cols = np.arange(int(np.ceil((5.0 - 1.8) / dx)), int((35.0 - 1.8) / dx) + 1)
    a, b, c = u[0][:, :, :, cols], u[1][:, :, :, 2 * cols], u[2][:, :, :, 4 * cols]
    for f in (0, 1, 4):
        e1 = np.abs(a[f] - b[f]).max()
        e2 = np.abs(b[f] - c[f]).max()
        print("fd4 field", f, "differences", e1, e2, "ratio", e1 / e2, flush=True)
        assert 13.0 < e1 / e2 < 19.0, (f, e1 / e2)


def test_fd4_rejections():
    p = CASES[0][1]
    with pytest.raises(KGError) as ei:
        LinesMethod(Nx=p["Nx"], Ny=p["Ny"], Nz=p["Nz"], xmin=p["xmin"], xmax=p["xmax"], scheme="pirk2",
                    operators="fd4")
    assert ei.value.status == capi.KG_ERR_UNSUPPORTED
    lm = LinesMethod(Nx=p["Nx"], Ny=p["Ny"], Nz=p["Nz"], xmin=p["xmin"], xmax=p["xmax"], scheme="rk4",
                     operators="fd4")
    with pytest.raises(KGError) as ei:
        lm.pirk2()
    assert ei.value.status == capi.KG_ERR_UNSUPPORTED
    lm.close()


def test_fd4_small_reference_workload_2000_steps():
    """The small workload as specified (Nr=256, Ntheta=32, Nphi=64, r in [1.8, 60],
    massless, Kerr-Schild M=1, RK4 at CFL 0.25, 4th-order stencils, Sommerfeld
    outer boundary, 2000 steps, phi = 1e-3 exp(-(r-15)^2/1.5^2) Y_2^0, Pi = 0):
    bit-identical to the oracle, so the per-field relative error is 0 against
    the 1e-12 bar."""
    from kg_amd.initial_data import small_workload_state
    print("fd4 small workload, Sommerfeld, 2000 steps", flush=True)
    sim = oracle.Sim(256, 33, 65, 2, 1.8, 60.0, "Kerr_Schild", 0.0, 1.0, 0.0, "sommerfeld", 0.25)
    sim.set_order(4)
    u0 = sim.fill(small_workload_state())
    sim.u[...] = u0
    gpu = LinesMethod(Nx=256, Ny=33, Nz=65, xmin=1.8, xmax=60.0, CFL=0.25, scheme="rk4", operators="fd4",
                      boundary_type_rmax="sommerfeld")
    gpu.set_u_soa(u0[None])
    for _ in range(4):
        sim.step("rk4", 500)
        gpu.step(500)
        print("  500 steps more", flush=True)
    got = gpu.u_soa()[0]
    assert np.isfinite(got).all() and np.abs(got).max() > 0
    assert max(_rel_err(got, sim.u)) <= 1e-12
    assert np.array_equal(got, sim.u)
    gpu.close()
