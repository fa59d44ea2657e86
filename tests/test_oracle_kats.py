"""Known-answer and invariant checks that pin the CPU oracle.

The reference ships no tests or golden data.  Each check below is fixed by
the mathematics or by the reference's own text, and is chosen so that a
plausible slip in the oracle (a dropped term, a wrong sign, index or
coefficient) fails it:
  - grid, dt: the formulas of mesh_refinement_lib.f90:109-132 and
    diagnostic_lib.f90:101-104 evaluated independently;
  - metric: closed-form Schwarzschild in Kerr-Schild coordinates; the
    reference's own inverse-metric identity (geometry.f90:409-492); the
    theta-ghost row copy (geometry.f90:371-377);
  - stencils: exactness on quadratics (the reference stencils themselves are
    pinned bit for bit in test_oracle_reference_stencils.py);
  - ghost fill: index tags against the maps of lines_method_lib.f90:507-597;
  - integrators: amplification polynomials, the PIRK recurrences written
    as matrices, observed orders of accuracy, symmetry invariants.
"""
import math

import numpy as np
import pytest

import oracle


# ----------------------------------------------------------------------------
# grid and dt
# ----------------------------------------------------------------------------
def test_grid_coordinates_and_dt():
    Nx, Ny, Nz, g, xmin, xmax = 256, 33, 65, 2, 1.8, 60.0
    s = oracle.Sim(Nx, Ny, Nz, g, xmin, xmax)
    dx = (xmax - xmin) / Nx
    dy = math.pi / (Ny - 1)
    dz = 2 * math.pi / (Nz - 1)
    assert s.dx == dx and s.dy == dy and s.dz == dz
    x = s.table("x")
    assert np.array_equal(x, xmin + np.arange(Nx + 2 * g + 1) * dx)
    y = s.table("y")
    assert np.array_equal(y, -dy / 2 + (np.arange(Ny + 2 * g + 1) - g) * dy)
    assert np.array_equal(s.table("z"), np.arange(Nz + 1) * dz)
    # theta is cell-centred: rows j = 0 and j = Ny straddle the poles
    assert y[g] < 0 < y[g + 1] and y[g + Ny - 1] < math.pi < y[g + Ny]
    expect = 0.25 * min(dx, math.exp(xmin) * dy, math.exp(xmin) * math.sin(dy / 2) * dz)
    assert s.dt == pytest.approx(expect, rel=1e-15)
    assert s.dt == pytest.approx(7.2856e-3, rel=1e-4)
    s2dt = oracle.reference_dt(2048, 129, 257, 2, 1.8, 200.0, 0.25)
    assert s2dt == pytest.approx(4.5552e-4, rel=1e-4)


# ----------------------------------------------------------------------------
# metric
# ----------------------------------------------------------------------------
def test_schwarzschild_kerr_schild_closed_form():
    M = 1.3
    s = oracle.Sim(40, 17, 9, 2, 1.8, 20.0, "Kerr_Schild", 0.0, M, 0.0)
    r = s.table("x")[None, :]
    th = s.table("y")[:, None]
    h = 1.0 + 2.0 * M / r
    rows = slice(0, s.nj - 1)  # the last row is the copied ghost row, checked below
    checks = {
        "alpha": 1.0 / np.sqrt(h),
        "beta_r": 2.0 * M / (r + 2.0 * M),
        "g_rr": 1.0 / h,
        "g_tt": 1.0 / r ** 2 + 0 * th,
        "g_pp": 1.0 / (r ** 2 * np.sin(th) ** 2),
        "sqrtg": r ** 2 * np.abs(np.sin(th)) * np.sqrt(h),
    }
    for name, exact in checks.items():
        got = s.table(name)[rows]
        ex = np.broadcast_to(exact, (s.nj, s.ni))[rows]
        np.testing.assert_allclose(got, ex, rtol=2e-13, atol=0, err_msg=name)
    assert np.all(s.table("g_rp") == 0.0)


@pytest.mark.parametrize("metric,a", [("Kerr_Schild", 0.0), ("Kerr_Schild", 0.7), ("Schw_EF", 0.0),
                                      ("MKS", 0.5), ("Minkowski", 0.0)])
def test_inverse_metric_identity(metric, a):
    """validate_inverse_metric (geometry.f90:409-492): g_dn g_up = delta and
    g_dn(i,j) gamma_up(j,k) = delta at every padded point, tolerance 1e-10."""
    xmin, xmax = (0.6, 3.0) if metric == "MKS" else (1.9, 15.0)
    s = oracle.Sim(30, 13, 11, 2, xmin, xmax, metric, a, 1.0, 0.0)
    gdn, gup = s.table("g_dn"), s.table("g_up")
    prod = np.einsum("abjn,bcjn->acjn", gdn, gup)
    eye = np.eye(4)[:, :, None, None]
    assert np.abs(prod - eye).max() < 1e-10
    gam = np.zeros((3, 3, s.nj, s.ni))
    gam[0, 0], gam[0, 2], gam[2, 0] = s.table("g_rr"), s.table("g_rp"), s.table("g_rp")
    gam[1, 1], gam[2, 2] = s.table("g_tt"), s.table("g_pp")
    prod3 = np.einsum("abjn,bcjn->acjn", gdn[1:, 1:], gam)
    assert np.abs(prod3 - np.eye(3)[:, :, None, None]).max() < 1e-10


def test_theta_ghost_row_is_a_copy():
    """boundary_theta_dn/up (geometry.f90:371-377,398-404): row j = Ny+g takes
    the values of row Ny+g-1 instead of its own theta."""
    s = oracle.Sim(24, 11, 9, 2, 1.8, 10.0, "Kerr_Schild", 0.4, 1.0, 0.2)
    for name in ("alpha", "beta_r", "dx_beta", "sqrtg", "g_rr", "g_rp", "g_tt", "g_pp"):
        t = s.table(name)
        assert np.array_equal(t[-1], t[-2]), name
        assert not np.array_equal(t[-2], t[-3]) or name in ("alpha", "beta_r", "dx_beta", "g_rr")


def test_dx_beta_is_second_order():
    errs = []
    for Nx in (160, 320, 640):
        s = oracle.Sim(Nx, 9, 9, 2, 1.8, 20.0, "Kerr_Schild", 0.0, 1.0, 0.0)
        r = s.table("x")
        exact = -2.0 / (r + 2.0) ** 2   # d/dr of beta^r = 2M/(r+2M), M = 1
        errs.append(np.abs(s.table("dx_beta")[4] - exact).max())
    assert 3.5 < errs[0] / errs[1] < 4.5 and 3.5 < errs[1] / errs[2] < 4.5, errs


# ----------------------------------------------------------------------------
# stencils: exact on quadratics
# ----------------------------------------------------------------------------
def test_stencils_exact_on_quadratics():
    ni, nj, nk = 14, 10, 9
    h = 0.25
    i = np.arange(ni) * h
    f = (1.5 - 0.7 * i + 0.3 * i ** 2)[None, None, :] * np.ones((nk, nj, 1))
    d = oracle.stencil("dx1", f, h)
    np.testing.assert_allclose(d, np.broadcast_to(-0.7 + 0.6 * i, f.shape), rtol=0, atol=1e-13)
    for bval in (1.0, -1.0, 0.0):
        b = np.full(f.shape, bval)
        da = oracle.stencil("advec_x", f, h, beta=b)
        np.testing.assert_allclose(da, np.broadcast_to(-0.7 + 0.6 * i, f.shape), rtol=0, atol=1e-13)
    jv = np.arange(nj) * h
    fy = (2.0 + 0.4 * jv - 0.9 * jv ** 2)[None, :, None] * np.ones((nk, 1, ni))
    dy = oracle.stencil("dx2", fy, h)
    np.testing.assert_allclose(dy[:, 1:, :], np.broadcast_to((0.4 - 1.8 * jv)[None, :, None], fy.shape)[:, 1:, :],
                               rtol=0, atol=1e-12)
    kv = np.arange(nk) * h
    fz = (0.5 + 1.1 * kv + 0.2 * kv ** 2)[:, None, None] * np.ones((1, nj, ni))
    dz = oracle.stencil("dx3", fz, h)
    np.testing.assert_allclose(dz[1:-1], np.broadcast_to((1.1 + 0.4 * kv)[:, None, None], fz.shape)[1:-1],
                               rtol=0, atol=1e-12)


# ----------------------------------------------------------------------------
# ghost fill
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("Nz", [17, 14])
def test_fill_source_map(Nz):
    """Every padded point after boundary(): the evolved point it is copied from
    (lines_method_lib.f90:516-596, order xL, xR, yL, yR, zL, zR)."""
    Nx, Ny, g = 10, 7, 2
    s = oracle.Sim(Nx, Ny, Nz, g, 1.8, 8.0)
    ni, nj, nk = s.ni, s.nj, s.nk
    flat = lambda ii, jj, kk: (kk * nj + jj) * ni + ii
    kt = lambda k: (k + Nz // 2) % (Nz - 1)

    def after_x(ii, jj, kk):
        if ii == ni - 1:
            return flat(ni - 2, jj, kk)  # outflow copy
        return flat(ii, jj, kk)  # i = -g: extrapolation at its own row (tag == own index)

    def after_y(ii, jj, kk):
        j = jj - g
        if j <= 0:
            return after_x(ii, (1 - j) + g, kt(kk))
        if j >= Ny:
            return after_x(ii, (2 * Ny - j - 1) + g, kt(kk))
        return after_x(ii, jj, kk)

    def final(ii, jj, kk):
        if kk == 0:
            return after_y(ii, jj, Nz - 1)
        if kk == Nz:
            return after_y(ii, jj, 1)
        return after_y(ii, jj, kk)

    src = s.fill_source_map()
    for kk in range(nk):
        for jj in range(nj):
            for ii in range(ni):
                assert src[kk, jj, ii] == final(ii, jj, kk), (ii, jj, kk)


def test_second_fill_changes_only_the_k_star_ghost_rows():
    """The y-fill at k* (k_target(k*) = 0) reads plane 0 before the z-fill
    refreshes it (lines_method_lib.f90:326-329), so boundary() is not
    idempotent: a second call changes exactly those theta-ghost rows."""
    Nx, Ny, Nz, g = 10, 7, 17, 2
    s = oracle.Sim(Nx, Ny, Nz, g, 1.8, 8.0)
    rng = np.random.default_rng(4)
    u1 = s.fill(rng.standard_normal(s.shape))
    u2 = s.fill(u1.copy())
    diff = np.argwhere(u1 != u2)
    kstar = Nz - 1 - Nz // 2
    assert len(diff) > 0
    assert set(diff[:, 1]) == {kstar}
    rows = set(diff[:, 2])
    assert rows <= set(range(0, g + 1)) | set(range(Ny + g, Ny + 2 * g + 1))


# ----------------------------------------------------------------------------
# integrators
# ----------------------------------------------------------------------------
LAM = np.array([-1.0, 0.5, -2.0, 0.3, -0.7])


def _decoupled():
    return oracle.GenericIntegrator(4, lambda v: LAM[:4, None] * v[:4], lambda v: LAM[4] * v[4])


@pytest.mark.parametrize("scheme,poly", [
    ("rk4", lambda z: 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24),
    ("rk3", lambda z: 1 + z + z ** 2 / 2 + z ** 3 / 6),
])
def test_rk_amplification_polynomials(scheme, poly):
    gi = _decoupled()
    gi.u[:] = 1.0
    dt = 0.13
    gi.step(scheme, dt, 3)
    np.testing.assert_allclose(gi.u[:, 0], poly(LAM * dt) ** 3, rtol=2e-15)


def _pirk_matrix(scheme, dt, a, b, c, d):
    """One PIRK step on phi' = a phi + b Pi, Pi' = c phi + d Pi as a matrix on
    (phi, Pi, stage Pi of the previous step[, second stage Pi]), written from
    lines_method_lib.f90:275-308 (pirk2) and :337-385 (pirk3)."""
    n = 3 if scheme == "pirk2" else 4
    E = np.eye(n)
    u1, u5, s5o = E[0], E[1], E[2]
    R1 = lambda p1, p5: a * p1 + b * p5
    R5 = lambda p1, p5: c * p1 + d * p5
    s1 = u1 + dt * R1(u1, u5)
    s5 = u5 + dt * R5(s1, s5o)          # rhs_pi(u_stage) with the old u_stage(5)
    if scheme == "pirk2":
        v1 = 0.5 * (u1 + s1 + dt * R1(s1, s5))
        v5 = 0.5 * (u5 + s5 + dt * R5(s1, s5))
        return np.array([v1, v5, s5])
    s25o = E[3]
    t1 = 0.75 * u1 + 0.25 * (s1 + dt * R1(s1, s5))
    t5 = 0.75 * u5 + 0.25 * (s5 + dt * R5(t1, s25o))
    v1 = (1 / 3) * u1 + (2 / 3) * (t1 + dt * R1(t1, t5))
    v5 = (1 / 3) * u5 + (2 / 3) * (t5 + dt * R5(t1, t5))
    return np.array([v1, v5, s5, t5])


@pytest.mark.parametrize("scheme", ["pirk2", "pirk3"])
def test_pirk_recurrence_with_stale_pi(scheme):
    a, b, c, d = 0.2, 1.0, -1.3, -0.4
    gi = oracle.GenericIntegrator(1, lambda v: np.array([a * v[0] + b * v[4], 0 * v[0], 0 * v[0], 0 * v[0]]),
                                  lambda v: c * v[0] + d * v[4])
    gi.u[0] = 0.8
    gi.u[4] = -0.3
    gi.stage_array(0)[4] = 0.05   # a nonzero "previous" u_stage(5)
    gi.stage_array(1)[4] = -0.02
    dt = 0.07
    x = np.array([0.8, -0.3, 0.05] + ([-0.02] if scheme == "pirk3" else []))
    M = _pirk_matrix(scheme, dt, a, b, c, d)
    for _ in range(6):
        gi.step(scheme, dt, 1)
        x = M @ x
        np.testing.assert_allclose([gi.u[0, 0], gi.u[4, 0], gi.stage_array(0)[4, 0]], x[:3], rtol=1e-13)
        if scheme == "pirk3":
            np.testing.assert_allclose(gi.stage_array(1)[4, 0], x[3], rtol=1e-13)


@pytest.mark.parametrize("scheme,order", [("rk4", 4), ("rk3", 3), ("pirk2", 1), ("pirk3", 1)])
def test_observed_order_on_oscillator(scheme, order):
    """phi' = Pi, Pi' = -phi.  RK4 and SSP-RK3 show their classical orders.
    The reference's PIRK2/PIRK3 evaluate the Pi right-hand side of stage 1 on
    the new stage values and reuse those in the last stage, which makes both
    first order on this coupled system; the oracle must show the same."""
    errs = []
    for n in (40, 80, 160):
        dt = 1.0 / n
        gi = oracle.GenericIntegrator(1, lambda v: np.array([v[4], 0 * v[0], 0 * v[0], 0 * v[0]]),
                                      lambda v: -v[0])
        gi.u[0] = 1.0
        gi.step(scheme, dt, n)
        errs.append(max(abs(gi.u[0, 0] - math.cos(1.0)), abs(gi.u[4, 0] + math.sin(1.0))))
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(abs(r - order) < 0.15 for r in rates), rates


# ----------------------------------------------------------------------------
# the Klein-Gordon operator
# ----------------------------------------------------------------------------
def test_rhs_of_uniform_phi_is_mass_term():
    s = oracle.Sim(16, 9, 11, 2, 1.8, 9.0, "Kerr_Schild", 0.3, 1.0, 0.4)
    v = np.zeros(s.shape)
    v[0] = 2.5
    R = s.rhs_all(v)
    assert np.all(R[:4] == 0.0)
    am2 = s.table("alpha") * (0.4 * 0.4)
    assert np.array_equal(R[4], np.broadcast_to(-(am2 * 2.5), R[4].shape))


def test_axisymmetric_data_stays_axisymmetric():
    from kg_amd.initial_data import shell_state
    s = oracle.Sim(20, 9, 13, 2, 1.8, 10.0, "Kerr_Schild", 0.0, 1.0, 0.2)
    s.u[...] = shell_state(20, 9, 13, 2, 1.8, 10.0, 5.0, 1.5, 1.0, {(2, 0): 1.0, (1, 0): 0.3})
    for scheme in ("pirk2", "rk4"):
        s.step(scheme, 5)
        u = s.u
        assert np.all(u == u[:, :1]), scheme   # every phi plane identical, bit for bit
        assert np.all(u[3] == 0.0)


def test_rk4_time_convergence_of_full_operator():
    """RK4 on the full Klein-Gordon operator (fixed spatial grid), starting
    from ghost-filled data: halving dt reduces the difference to a fine-dt
    solution by ~16.  (Unfilled initial ghosts, inconsistent with the fill
    maps, make the first step's error O(dt) instead.)"""
    from kg_amd.initial_data import shell_state
    Nx, Ny, Nz = 24, 9, 13
    u0 = shell_state(Nx, Ny, Nz, 2, 1.8, 12.0, 6.0, 1.5, 1.0, {(1, 0): 1.0, (2, 1): 0.5})
    T = 0.4

    def run(nsteps):
        s = oracle.Sim(Nx, Ny, Nz, 2, 1.8, 12.0, "Kerr_Schild", 0.0, 1.0, 0.1, "outflow", 0.25, T / nsteps)
        s.u[...] = s.fill(u0.copy())
        s.step("rk4", nsteps)
        return s.u.copy()

    ref = run(512)
    e = [np.abs(run(n) - ref).max() for n in (16, 32)]
    assert 12.0 < e[0] / e[1] < 20.0, e


# ----------------------------------------------------------------------------
# the fd4 operator set (finite_differences.f90:124-151,241-268,340-367,922-957)
# ----------------------------------------------------------------------------
def test_fd4_stencils_exact_on_quartics():
    """Every row of x_4, y_4, z_4 and advec_x_4 (one-sided, biased, upwinded
    and central) differentiates a quartic exactly (up to rounding)."""
    ni, nj, nk = 15, 12, 11
    h = 0.25

    def quartic(t):
        return 1.5 - 0.7 * t + 0.3 * t ** 2 - 0.2 * t ** 3 + 0.05 * t ** 4, \
            -0.7 + 0.6 * t - 0.6 * t ** 2 + 0.2 * t ** 3

    i = np.arange(ni) * h
    p, dp = quartic(i)
    f = p[None, None, :] * np.ones((nk, nj, 1))
    want = np.broadcast_to(dp, f.shape)
    np.testing.assert_allclose(oracle.stencil("dx1_4", f, h), want, rtol=0, atol=1e-12)
    for bval in (1.0, -1.0, 0.0):
        da = oracle.stencil("advec_x_4", f, h, beta=np.full(f.shape, bval))
        np.testing.assert_allclose(da, want, rtol=0, atol=1e-12)
    jv = np.arange(nj) * h
    p, dp = quartic(jv)
    fy = p[None, :, None] * np.ones((nk, 1, ni))
    np.testing.assert_allclose(oracle.stencil("dx2_4", fy, h), np.broadcast_to(dp[None, :, None], fy.shape),
                               rtol=0, atol=1e-12)
    kv = np.arange(nk) * h
    p, dp = quartic(kv)
    fz = p[:, None, None] * np.ones((1, nj, ni))
    np.testing.assert_allclose(oracle.stencil("dx3_4", fz, h), np.broadcast_to(dp[:, None, None], fz.shape),
                               rtol=0, atol=1e-12)
    # a quintic is not differentiated exactly: the pin has teeth
    f5 = (i ** 5)[None, None, :] * np.ones((nk, nj, 1))
    assert np.abs(oracle.stencil("dx1_4", f5, h) - 5 * i ** 4).max() > 1e-6


def test_fd4_dx_beta_is_fourth_order():
    """With the fd4 set the metric's d_r beta^r is x_4 of beta^r: its error
    against 2M/(r+2M)' = -2M/(r+2M)^2 falls ~16x per halving of dr."""
    errs = []
    for n in (64, 128, 256):
        s = oracle.Sim(n, 5, 5, 2, 1.8, 9.0, "Kerr_Schild", 0.0, 1.0, 0.0)
        s.set_order(4)
        x = s.table("x")
        exact = -2.0 / (x + 2.0) ** 2
        errs.append(np.abs(s.table("dx_beta")[3] - exact).max())
    assert 13.0 < errs[0] / errs[1] < 19.0 and 14.0 < errs[1] / errs[2] < 18.0, errs


def test_fd4_rhs_uniform_phi_and_axisymmetry():
    from kg_amd.initial_data import shell_state
    s = oracle.Sim(16, 9, 11, 2, 1.8, 9.0, "Kerr_Schild", 0.3, 1.0, 0.4)
    s.set_order(4)
    v = np.zeros(s.shape)
    v[0] = 2.5
    R = s.rhs_all(v)
    assert np.all(R[:4] == 0.0)
    assert np.array_equal(R[4], np.broadcast_to(-(s.table("alpha") * (0.4 * 0.4) * 2.5), R[4].shape))
    s = oracle.Sim(20, 9, 13, 2, 1.8, 10.0, "Kerr_Schild", 0.0, 1.0, 0.2)
    s.set_order(4)
    s.u[...] = shell_state(20, 9, 13, 2, 1.8, 10.0, 5.0, 1.5, 1.0, {(2, 0): 1.0, (1, 0): 0.3})
    s.step("rk4", 5)
    # the one-sided phi rows of z_4 (-25c + 48c - 36c + 16c - 3c) leave the
    # rounding of a constant, so the planes agree to rounding, not bit for bit
    scale = np.abs(s.u).max()
    assert np.abs(s.u - s.u[:, :1]).max() <= 1e-13 * scale
    assert np.abs(s.u[3]).max() <= 1e-13 * scale


def test_fd4_rk4_self_convergence_is_fourth_order():
    """A spherically symmetric massive pulse on Schwarzschild-KS (only the
    radial operators and RK4 carry error; dt proportional to dr): with the
    fd4 set the differences between three resolutions fall by ~2^4."""
    from kg_amd.initial_data import shell_state

    def run(Nx, T=3.0, xmin=1.8, xmax=21.8):
        dt = 0.25 * (xmax - xmin) / Nx
        s = oracle.Sim(Nx, 5, 5, 2, xmin, xmax, "Kerr_Schild", 0.0, 1.0, 0.2, "outflow", 0.25, dt)
        s.set_order(4)
        s.u[...] = s.fill(shell_state(Nx, 5, 5, 2, xmin, xmax, 10.0, 1.5, 1.0, {(0, 0): 1.0}))
        n = int(round(T / dt))
        s.step("rk4", n)
        return s.u.copy()

    u = [run(N) for N in (80, 160, 320)]
    dx = 20.0 / 80
    cols = np.arange(int(np.ceil((4.0 - 1.8) / dx)), int((18.0 - 1.8) / dx) + 1)
    a, b, c = u[0][:, :, :, cols], u[1][:, :, :, 2 * cols], u[2][:, :, :, 4 * cols]
    for f in (0, 1, 4):
        q = np.abs(a[f] - b[f]).max() / np.abs(b[f] - c[f]).max()
        assert 13.0 < q < 19.0, (f, q)


def test_oracle_results_do_not_depend_on_the_thread_count():
    """The oracle's point loops run on OpenMP threads; every scheme and both
    operator sets give the same bits on 1 thread as on all of them."""
    import os
    import subprocess
    import sys
    code = ("import hashlib, oracle\n"
            "from kg_amd.initial_data import shell_state, random_lm_coeffs\n"
            "h = hashlib.sha256()\n"
            "for order, scheme in ((2, 'pirk2'), (2, 'rk4'), (4, 'rk4'), (4, 'rk3'), (2, 'pirk3')):\n"
            "    s = oracle.Sim(30, 9, 13, 2, 1.8, 12.0, 'Kerr_Schild', 0.5, 1.0, 0.2, 'inflow', 0.25)\n"
            "    s.set_order(order)\n"
            "    s.u[...] = shell_state(30, 9, 13, 2, 1.8, 12.0, 6.0, 1.5, 1.0, random_lm_coeffs(3, 3, 0.3))\n"
            "    s.step(scheme, 5)\n"
            "    h.update(s.u.tobytes())\n"
            "print(h.hexdigest())\n")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = []
    for nt in ("1", "4"):
        env = dict(os.environ, OMP_NUM_THREADS=nt, PYTHONPATH=root)
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, cwd=root,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        out.append(r.stdout.strip())
    assert out[0] == out[1]
