"""Pins of the oracle's right-hand side and metric that mathematics fixes.

The oracle restates the reference's RHS assembly (lines_method_lib.f90:671-753)
and its Kerr-Schild tables (geometry.f90:112-160,286-342); the HIP engine
reproduces the oracle bit for bit, so these checks are what ties both to the
Klein-Gordon equation itself.

1. Metric.  The reference's Kerr-Schild g_{mu nu} equals the textbook Kerr
   metric in Boyer-Lindquist coordinates pulled back by the ingoing
   Kerr-Schild map dt_KS = dt_BL + 2Mr/Delta dr, dphi_KS = dphi_BL + a/Delta dr;
   the 3+1 tables (alpha, beta^r, gamma^{ij}, sqrt(gamma)) equal their closed
   forms with q = 2Mr/rho^2: alpha = (1+q)^(-1/2), beta^r = q/(1+q),
   gamma^{rr} = Delta/rho^2 + q^2/(1+q), gamma^{r phi} = a/rho^2,
   gamma^{theta theta} = 1/rho^2, gamma^{phi phi} = 1/(rho^2 sin^2 theta),
   sqrt(gamma) = rho^2 |sin theta| sqrt(1+q), and beta^theta = beta^phi = 0.
2. Operator (manufactured solution).  For any smooth phi(t, r, theta, phi),
   data phi, psi_i = d_i phi, Pi = (d_t phi - beta^r psi_r)/alpha at t = 0 give
   the exact time derivatives R_phi = d_t phi, R_psi_i = d_i d_t phi and
   R_Pi = d_t Pi + alpha (box phi - m^2 phi), where box phi =
   (-g)^(-1/2) d_mu((-g)^(1/2) g^{mu nu} d_nu phi) is taken symbolically
   (sympy) from the closed-form inverse 4-metric (g^{tt} = -(1+q),
   g^{tr} = q, g^{rr} = Delta/rho^2, g^{r phi} = a/rho^2, g^{theta theta} =
   1/rho^2, g^{phi phi} = 1/(rho^2 sin^2 theta), sqrt(-g) = rho^2 sin theta;
   itself checked against the pulled-back metric).  The discrete RHS minus
   these converges at the operator set's order (ratio 4 per halving for the
   2nd-order set, 16 for fd4) away from the edges and poles, for a spinning
   hole with a field mass.
3. Static solutions of box phi = 0 (checked symbolically): ln(1 - 2M/r) and
   (r - M) sin(theta) cos(phi) on Schwarzschild, ln((r - r+)/(r - r-)) on
   Kerr: the RHS of their data converges to zero at the same orders.
4. Mutation checks: each plausible slip the oracle can plant (a dropped
   term, a swapped coefficient, a wrong metric component) fails (2) or (1).
"""
import math

import numpy as np
import pytest
import sympy as sp

import oracle

T_, R_, TH_, PH_ = sp.symbols("t r theta phi", real=True)


# ---------------------------------------------------------------------------
# metric
# ---------------------------------------------------------------------------
def kerr_bl_pullback(r, th, a, M):
    """Kerr in Boyer-Lindquist coordinates pulled back to ingoing Kerr-Schild:
    g_KS = J^T g_BL J with J = d(t_BL, r, theta, phi_BL)/d(t, r, theta, phi)."""
    rho2 = r ** 2 + a * a * np.cos(th) ** 2
    Delta = r ** 2 - 2 * M * r + a * a
    s2 = np.sin(th) ** 2
    g = np.zeros((4, 4) + np.shape(r))
    g[0, 0] = -(1 - 2 * M * r / rho2)
    g[0, 3] = g[3, 0] = -2 * M * a * r * s2 / rho2
    g[1, 1] = rho2 / Delta
    g[2, 2] = rho2
    g[3, 3] = (r ** 2 + a * a + 2 * M * a * a * r * s2 / rho2) * s2
    J = np.zeros_like(g)
    for i in range(4):
        J[i, i] = 1.0
    J[0, 1] = -2 * M * r / Delta
    J[3, 1] = -a / Delta
    return np.einsum("ai...,ab...,bj...->ij...", J, g, J)


def closed_form_31(r, th, a, M):
    rho2 = r ** 2 + a * a * np.cos(th) ** 2
    Delta = r ** 2 - 2 * M * r + a * a
    q = 2 * M * r / rho2
    return {
        "alpha": 1.0 / np.sqrt(1 + q),
        "beta_r": q / (1 + q),
        "g_rr": Delta / rho2 + q * q / (1 + q),
        "g_rp": a / rho2,
        "g_tt": 1.0 / rho2,
        "g_pp": 1.0 / (rho2 * np.sin(th) ** 2),
        "sqrtg": rho2 * np.abs(np.sin(th)) * np.sqrt(1 + q),
    }


def closed_form_inverse(r, th, a, M):
    rho2 = r ** 2 + a * a * np.cos(th) ** 2
    Delta = r ** 2 - 2 * M * r + a * a
    q = 2 * M * r / rho2
    gi = np.zeros((4, 4) + np.shape(r))
    gi[0, 0] = -(1 + q)
    gi[0, 1] = gi[1, 0] = q
    gi[1, 1] = Delta / rho2
    gi[1, 3] = gi[3, 1] = a / rho2
    gi[2, 2] = 1 / rho2
    gi[3, 3] = 1 / (rho2 * np.sin(th) ** 2)
    return gi


def _grid(sim):
    x, y = sim.table("x"), sim.table("y")
    R, TH = np.meshgrid(x, y)  # [nj][ni]
    return R, TH


SPINS = [(0.0, 1.0), (0.7, 1.0), (-0.5, 1.7), (0.95, 0.8)]


@pytest.mark.parametrize("a,M", SPINS)
def test_kerr_schild_metric_is_kerr_pulled_back_from_boyer_lindquist(a, M):
    sim = oracle.Sim(24, 13, 9, 2, 2.6 * M, 15.0, "Kerr_Schild", a, M, 0.0, "outflow", 0.25)
    R, TH = _grid(sim)
    want = kerr_bl_pullback(R, TH, a, M)
    got = sim.table("g_dn")
    rows = slice(0, sim.nj - 1)  # the last theta-ghost row is a copy of the one before (geometry.f90:371-377)
    scale = np.abs(want[..., rows, :]).max()
    assert np.abs(got[..., rows, :] - want[..., rows, :]).max() <= 1e-13 * scale


@pytest.mark.parametrize("a,M", SPINS)
def test_closed_form_inverse_is_the_inverse(a, M):
    """The closed-form g^{mu nu} used for box phi below inverts the pulled-back
    metric, and det g = -rho^4 sin^2 theta."""
    r = np.linspace(2.6 * M, 15.0, 17)[None, :]
    th = np.linspace(0.05, math.pi - 0.05, 11)[:, None]
    R, TH = np.broadcast_arrays(r, th)
    g = kerr_bl_pullback(R, TH, a, M)
    gi = closed_form_inverse(R, TH, a, M)
    prod = np.einsum("ab...,bc...->ac...", gi, g)
    assert np.abs(prod - np.eye(4)[:, :, None, None]).max() < 1e-12
    det = np.linalg.det(np.moveaxis(g, (0, 1), (-2, -1)))
    rho2 = R ** 2 + a * a * np.cos(TH) ** 2
    assert np.abs(det / (-(rho2 ** 2) * np.sin(TH) ** 2) - 1).max() < 1e-12


@pytest.mark.parametrize("a,M", SPINS)
def test_three_plus_one_tables_closed_form(a, M):
    sim = oracle.Sim(24, 13, 9, 2, 2.6 * M, 15.0, "Kerr_Schild", a, M, 0.0, "outflow", 0.25)
    R, TH = _grid(sim)
    rows = slice(0, sim.nj - 1)
    cf = closed_form_31(R, TH, a, M)
    for name, want in cf.items():
        got = sim.table(name)
        scale = np.abs(want[rows]).max()
        if scale == 0.0:  # gamma^{r phi} of a non-spinning hole
            assert np.abs(got[rows]).max() == 0.0, name
            continue
        rel = np.abs(got[rows] - want[rows]).max() / scale
        assert rel <= 2e-13, (name, rel)
    gup = sim.table("g_up")
    # beta^theta = alpha^2 g^{t theta}, beta^phi = alpha^2 g^{t phi}: both vanish
    assert np.abs(gup[0, 2][rows]).max() == 0.0
    assert np.abs(gup[0, 3][rows]).max() <= 1e-15 * np.abs(gup[0, 0][rows]).max()


# ---------------------------------------------------------------------------
# the operator
# ---------------------------------------------------------------------------
def _box(f, a, M):
    rho2 = R_ ** 2 + a ** 2 * sp.cos(TH_) ** 2
    Delta = R_ ** 2 - 2 * M * R_ + a ** 2
    q = 2 * M * R_ / rho2
    gi = {(0, 0): -(1 + q), (0, 1): q, (1, 0): q, (1, 1): Delta / rho2, (1, 3): a / rho2, (3, 1): a / rho2,
          (2, 2): 1 / rho2, (3, 3): 1 / (rho2 * sp.sin(TH_) ** 2)}
    sq = rho2 * sp.sin(TH_)
    X = (T_, R_, TH_, PH_)
    box = 0
    for mu in range(4):
        inner = sum(sq * gi[(mu, nu)] * sp.diff(f, X[nu]) for nu in range(4) if (mu, nu) in gi)
        box += sp.diff(inner, X[mu])
    return box / sq, q


def exact_fields(A, B, C, a, M, m):
    """phi = A + t B + t^2 C (functions of r, theta, phi): the state at t = 0 and
    the exact right-hand side, as numpy functions of (r, theta, phi)."""
    f = A + T_ * B + T_ ** 2 * C
    box, q = _box(f, a, M)
    alpha = 1 / sp.sqrt(1 + q)
    beta = q / (1 + q)
    Pi = (B - beta * sp.diff(A, R_)) / alpha
    r_pi = (2 * C - beta * sp.diff(B, R_)) / alpha + alpha * (box.subs(T_, 0) - m ** 2 * A)
    state = [A, sp.diff(A, R_), sp.diff(A, TH_), sp.diff(A, PH_), Pi]
    rhs = [B, sp.diff(B, R_), sp.diff(B, TH_), sp.diff(B, PH_), r_pi]
    lam = lambda e: sp.lambdify((R_, TH_, PH_), e, "numpy")
    return [lam(e) for e in state], [lam(e) for e in rhs]


def _residual(order, Nr, case, state, rhs, plant=0):
    """max |RHS_oracle - RHS_exact| / max |RHS_exact| per field on the
    comparison region, with (Ntheta, Nphi) = (Nr/2, Nr)."""
    Ny, Nz = Nr // 2 + 1, Nr + 1
    sim = oracle.Sim(Nr, Ny, Nz, 2, case["xmin"], case["xmax"], case["metric"], case["a"], case["M"], case["m"],
                     "outflow", 0.25)
    sim.set_order(order)
    if plant:
        sim.plant(plant)
    x, y, z = sim.table("x"), sim.table("y"), sim.table("z")
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    v = np.stack([np.broadcast_to(fn(X, Y, Z), X.shape) for fn in state]).astype(np.float64)
    got = sim.rhs_all(v)
    want = np.stack([np.broadcast_to(fn(X, Y, Z), X.shape) for fn in rhs]).astype(np.float64)
    sim.close()
    ii = (x >= case["r_lo"]) & (x <= case["r_hi"])
    jj = (y >= 0.5) & (y <= math.pi - 0.5)
    kk = np.arange(len(z)) < Nz - 1  # k = Nz is the periodic duplicate plane the fill overwrites
    sel = np.ix_(range(5), kk, jj, ii)
    d, w = (got - want)[sel], want[sel]
    absolute = [float(np.abs(d[c]).max()) for c in range(5)]
    wmax = [float(np.abs(w[c]).max()) for c in range(5)]
    rel = [absolute[c] / wmax[c] if wmax[c] > 0 else absolute[c] for c in range(5)]
    return rel, wmax, absolute


CASES = {
    "kerr_a0.7_m0.3": dict(metric="Kerr_Schild", a=0.7, M=1.0, m=0.3, xmin=3.0, xmax=15.0, r_lo=5.0, r_hi=13.0),
    "schwarzschild_m0.1": dict(metric="Kerr_Schild", a=0.0, M=1.0, m=0.1, xmin=3.0, xmax=15.0, r_lo=5.0, r_hi=13.0),
    "schw_ef_m0.2": dict(metric="Schw_EF", a=0.0, M=1.0, m=0.2, xmin=3.0, xmax=15.0, r_lo=5.0, r_hi=13.0),
    "kerr_a-0.5_M1.5": dict(metric="Kerr_Schild", a=-0.5, M=1.5, m=0.0, xmin=4.0, xmax=16.0, r_lo=6.0, r_hi=14.0),
}


def _manufactured(case):
    A = sp.exp(-(R_ - 9) ** 2 / 9) * (1 + sp.Rational(3, 10) * sp.sin(TH_) * sp.cos(PH_)
                                      + sp.Rational(1, 5) * sp.cos(TH_) ** 2 * sp.sin(2 * PH_))
    B = sp.exp(-(R_ - 8) ** 2 / 6) * (sp.Rational(1, 2) + sp.Rational(1, 4) * sp.sin(TH_) * sp.sin(PH_)
                                      + sp.Rational(1, 5) * sp.cos(TH_))
    C = sp.exp(-(R_ - 10) ** 2 / 8) * (sp.Rational(1, 3) * sp.cos(TH_)
                                       + sp.Rational(1, 7) * sp.sin(TH_) ** 2 * sp.cos(2 * PH_))
    return exact_fields(A, B, C, case["a"], case["M"], case["m"])


ORDERS = {2: (3.5, 4.5), 4: (13.0, 19.0)}


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("name", list(CASES))
def test_rhs_converges_to_the_klein_gordon_equation(order, name):
    case = CASES[name]
    state, rhs = _manufactured(case)
    errs = [_residual(order, N, case, state, rhs)[0] for N in (24, 48, 96)]
    lo, hi = ORDERS[order]
    # R_phi = alpha Pi + beta psi_r is d_t phi exactly (up to rounding)
    assert max(e[0] for e in errs) < 1e-14
    for c in range(1, 5):
        for i in range(2):
            ratio = errs[i][c] / errs[i + 1][c]
            assert lo < ratio < hi, (name, order, "field", c, [e[c] for e in errs])


STATIC = {
    # (solution, a, M, xmin, xmax, r_lo, r_hi)
    "ln(1-2M/r), Schwarzschild": (sp.log(1 - 2 / R_), 0.0, 1.0, 3.0, 15.0, 5.0, 13.0),
    "(r-M) sin(theta) cos(phi), Schwarzschild": ((R_ - 1) * sp.sin(TH_) * sp.cos(PH_), 0.0, 1.0, 3.0, 15.0, 5.0, 13.0),
    "ln((r-r+)/(r-r-)), Kerr a=0.7": (
        sp.log((R_ - (1 + sp.sqrt(sp.Rational(51, 100)))) / (R_ - (1 - sp.sqrt(sp.Rational(51, 100))))),
        sp.Rational(7, 10), 1.0, 3.0, 15.0, 5.0, 13.0),
}


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("name", list(STATIC))
def test_static_solutions_have_converging_zero_rhs(order, name):
    sol, a, M, xmin, xmax, r_lo, r_hi = STATIC[name]
    box, _ = _box(sol, a, M)
    assert sp.simplify(box) == 0  # a static massless solution
    case = dict(metric="Kerr_Schild", a=float(a), M=M, m=0.0, xmin=xmin, xmax=xmax, r_lo=r_lo, r_hi=r_hi)
    state, rhs = exact_fields(sol, sp.Integer(0), sp.Integer(0), a, M, 0)
    errs = []
    for N in (24, 48, 96):
        _, w, e = _residual(order, N, case, state, rhs)
        assert max(w) < 1e-12  # the exact right-hand side vanishes
        errs.append(e)
    # the exact RHS is zero, so these are absolute residuals
    lo, hi = ORDERS[order]
    for c in range(1, 5):
        if errs[0][c] < 1e-10:  # a field whose data makes this derivative vanish identically
            continue
        for i in range(2):
            ratio = errs[i][c] / errs[i + 1][c]
            assert lo < ratio < hi, (name, order, "field", c, [e[c] for e in errs])


@pytest.mark.parametrize("plant", list(oracle.PLANTS))
def test_planted_mistakes_are_caught(plant):
    """Each planted slip leaves some field of the manufactured-solution residual
    at O(1e-2) and not converging, with either operator set (and the wrong g_dn
    component also fails the closed-form tables)."""
    code = oracle.PLANTS[plant]
    case = CASES["kerr_a0.7_m0.3"]
    state, rhs = _manufactured(case)
    for order in (2, 4):
        e1 = _residual(order, 48, case, state, rhs, plant=code)[0]
        e2 = _residual(order, 96, case, state, rhs, plant=code)[0]
        caught = [c for c in range(5) if e2[c] > 1e-3 and e1[c] / e2[c] < 2.0]
        assert caught, (plant, order, e1, e2)
    if plant == "gdn13_no_sin2":
        sim = oracle.Sim(24, 13, 9, 2, 2.6, 15.0, "Kerr_Schild", 0.7, 1.0, 0.0, "outflow", 0.25)
        sim.plant(code)
        R, TH = _grid(sim)
        rows = slice(0, sim.nj - 1)
        assert np.abs(sim.table("g_dn") - kerr_bl_pullback(R, TH, 0.7, 1.0))[..., rows, :].max() > 1e-2
        want = closed_form_31(R, TH, 0.7, 1.0)["g_rp"]
        assert np.abs(sim.table("g_rp")[rows] - want[rows]).max() > 1e-3
