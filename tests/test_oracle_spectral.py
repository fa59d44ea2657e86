"""Pins of the spectral-projection oracle (oracle/kg_spectral_oracle.c).

The reference takes P_l^m, j_l and the zeros of J_{l+1/2} from GSL
(tools/FFT_sph.f90:22-34,290-297; GSL is linked unversioned from
main/Makefile:8,27 and is absent here).  Its own tests pin nothing at this
boundary (test_orthonormality, FFT_sph.f90:338-384, is never called), so each
function is pinned here against closed forms (tools/harmonics.f90:55-65 for
Y_lm with l <= 2), scipy's independent implementations, and the mathematics:
orthogonality, quadrature convergence rates and exactness of trilinear
interpolation on trilinear functions.
"""
import math

import numpy as np
import pytest
from scipy.optimize import brentq
from scipy.special import lpmv, spherical_jn

import oracle


# ------------------------------------------------------------------ P_l^m
def test_plm_matches_scipy_with_condon_shortley_phase():
    xs = np.linspace(-0.999, 0.999, 101)
    for l in range(0, 13):
        for m in range(0, l + 1):
            ours = np.array([oracle.plm(l, m, x) for x in xs])
            ref = lpmv(m, l, xs)
            scale = max(np.abs(ref).max(), 1e-300)
            assert np.abs(ours - ref).max() <= 1e-13 * scale, (l, m)


def test_plm_closed_forms():
    for x in (-0.7, -0.2, 0.3, 0.9):
        s = math.sqrt(1 - x * x)
        assert oracle.plm(1, 1, x) == pytest.approx(-s, rel=1e-15)        # Condon-Shortley sign
        assert oracle.plm(2, 0, x) == pytest.approx(0.5 * (3 * x * x - 1), rel=1e-14, abs=1e-16)
        assert oracle.plm(2, 1, x) == pytest.approx(-3 * x * s, rel=1e-14)
        assert oracle.plm(2, 2, x) == pytest.approx(3 * (1 - x * x), rel=1e-14)
        assert oracle.plm(3, 3, x) == pytest.approx(-15 * s ** 3, rel=1e-14)


# ------------------------------------------------------------------ j_l
def test_sph_jl_closed_forms():
    for x in (0.3, 1.0, 2.5, 7.0, 31.0):
        s, c = math.sin(x), math.cos(x)
        assert oracle.sph_jl(0, x) == pytest.approx(s / x, rel=1e-15)
        assert oracle.sph_jl(1, x) == pytest.approx(s / x ** 2 - c / x, rel=1e-12, abs=1e-17)
        assert oracle.sph_jl(2, x) == pytest.approx((3 / x ** 2 - 1) * s / x - 3 * c / x ** 2, rel=1e-11, abs=1e-15)
    # small x, where the closed forms cancel: the series x^l/(2l+1)!! (1 - x^2/(2(2l+3)) + ...)
    for x in (1e-2, 1e-4):
        assert oracle.sph_jl(2, x) == pytest.approx(x * x / 15 * (1 - x * x / 14 + x ** 4 / 504), rel=1e-13)
        assert oracle.sph_jl(5, x) == pytest.approx(x ** 5 / 10395 * (1 - x * x / 26), rel=1e-12)
    assert oracle.sph_jl(0, 0.0) == 1.0 and oracle.sph_jl(3, 0.0) == 0.0


def test_sph_jl_matches_scipy():
    xs = np.concatenate([np.geomspace(1e-6, 1.0, 40), np.linspace(1.0, 250.0, 400)])
    for l in range(0, 11):
        ours = np.array([oracle.sph_jl(l, x) for x in xs])
        ref = spherical_jn(l, xs)
        # absolute to the function's scale (|j_l| <= 1), relative where it is tiny
        err = np.abs(ours - ref) / np.maximum(np.abs(ref), 1e-3)
        assert err.max() < 1e-11, (l, xs[err.argmax()], err.max())


def test_sph_jl_downward_recurrence_near_zeros_of_j0():
    """x = n*pi below the order: Miller's recurrence normalised by j_0 alone would
    divide by sin(x) ~ 1e-16 there."""
    for l in (4, 7, 10):
        for x in (math.pi, 2 * math.pi, 3 * math.pi, 3 * math.pi + 1e-9):
            if x < l:
                ref = spherical_jn(l, x)
                assert oracle.sph_jl(l, x) == pytest.approx(ref, rel=1e-12), (l, x)


def test_jl_zeros():
    for n in range(1, 8):
        assert oracle.jl_zero(0, n) == pytest.approx(n * math.pi, rel=1e-15)
    for l in range(1, 9):
        zs = [oracle.jl_zero(l, n) for n in range(1, 7)]
        for n, z in enumerate(zs, start=1):
            ref = brentq(lambda t: spherical_jn(l, t), z - 0.3, z + 0.3, xtol=1e-15)
            assert z == pytest.approx(ref, rel=1e-13), (l, n)
        # interlacing of the zeros of consecutive orders
        zn = [oracle.jl_zero(l - 1, n) for n in range(1, 8)]
        for n in range(6):
            assert zn[n] < zs[n] < zn[n + 1]


# ------------------------------------------------------------------ Y_lm
def _ylm_closed(l, m, th, ph):
    """tools/harmonics.f90:55-65"""
    pi = math.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    e = np.exp(1j * P)
    table = {
        (0, 0): 0.5 / math.sqrt(pi) * np.ones_like(e),
        (1, 1): -math.sqrt(3 / (8 * pi)) * e * np.sin(T),
        (1, 0): math.sqrt(3 / (4 * pi)) * np.cos(T) + 0j,
        (2, 2): math.sqrt(15 / (32 * pi)) * e ** 2 * np.sin(T) ** 2,
        (2, 1): -math.sqrt(15 / (8 * pi)) * e * np.sin(T) * np.cos(T),
        (2, 0): math.sqrt(5 / (16 * pi)) * (3 * np.cos(T) ** 2 - 1) + 0j,
    }
    if m >= 0:
        return table[(l, m)]
    return (-1) ** m * np.conj(table[(l, -m)])


def test_ylm_closed_forms_l_le_2():
    th = np.linspace(0.05, math.pi - 0.05, 23)
    ph = np.linspace(0.0, 2 * math.pi, 17)
    Y = oracle.sph_harmonics(2, th, ph)
    for l in range(3):
        for m in range(-l, l + 1):
            ref = _ylm_closed(l, m, th, ph)
            assert np.abs(Y[l * l + m + l] - ref).max() < 1e-14, (l, m)


def test_ylm_matches_scipy_up_to_l8():
    from scipy.special import sph_harm_y
    th = np.linspace(0.02, math.pi - 0.02, 31)
    ph = np.linspace(0.0, 2 * math.pi, 21)
    Y = oracle.sph_harmonics(8, th, ph)
    T, P = np.meshgrid(th, ph, indexing="ij")
    for l in range(9):
        for m in range(-l, l + 1):
            assert np.abs(Y[l * l + m + l] - sph_harm_y(l, m, T, P)).max() < 1e-13, (l, m)


def test_ylm_pole_rule():
    """|cos theta| > 0.9999: P_l^m := sqrt((2l+1)/4pi) for m = 0, 0 otherwise, so
    Y_l0 = (2l+1)/(4pi) at the poles (FFT_sph.f90:307-315)."""
    th = np.array([0.0, 0.005, 0.02, math.pi - 0.005, math.pi])
    ph = np.linspace(0, 2 * math.pi, 9)
    Y = oracle.sph_harmonics(4, th, ph)
    for l in range(5):
        for m in range(-l, l + 1):
            y = Y[l * l + m + l]
            for i in (0, 1, 3, 4):        # |cos| > 0.9999
                want = (2 * l + 1) / (4 * math.pi) if m == 0 else 0.0
                assert np.abs(y[i] - want).max() < 1e-15, (l, m, i)
            # theta = 0.02: cos = 0.9998 < 0.9999, the regular branch
            ref = np.array([oracle.plm(l, abs(m), math.cos(0.02))])
            assert abs(abs(y[2, 0]) - abs(ref[0]) * _norm(l, abs(m))) < 1e-14


def _norm(l, m):
    return math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - m) / math.factorial(l + m))


def _sph_grid(Ny, Nz):
    dth = math.pi / (Ny - 1)
    dph = 2 * math.pi / (Nz - 1)
    th = np.arange(Ny) * dth
    ph = np.arange(Nz) * dph
    return th, ph, dth, dph


def test_project_sh_recovers_coefficients_at_second_order():
    """f = sum c_lm Y_lm (l <= 4) projected back: the phi sum over a full period is
    exact, the theta trapezoid converges at O(dtheta^2)."""
    rng = np.random.default_rng(7)
    lmax = 4
    c = rng.normal(size=(lmax + 1) ** 2) + 1j * rng.normal(size=(lmax + 1) ** 2)
    errs = []
    for Ny in (33, 65, 129):
        th, ph, dth, dph = _sph_grid(Ny, 2 * (Ny - 1) + 1)
        Y = oracle.sph_harmonics(lmax, th, ph)
        f = np.tensordot(c, Y, axes=(0, 0))
        flm = oracle.project_sh(f, Y, lmax, th, dth, dph)
        errs.append(np.abs(flm - c).max())
    assert errs[2] < 2e-3
    assert 3.5 < errs[0] / errs[1] < 4.5 and 3.5 < errs[1] / errs[2] < 4.5, errs


def test_project_sh_m_orthogonality_is_exact():
    """Different m are orthogonal to rounding: the trapezoid with both duplicated
    phi end points at half weight is the exact discrete Fourier sum."""
    th, ph, dth, dph = _sph_grid(17, 33)
    Y = oracle.sph_harmonics(3, th, ph)
    for (l, m) in [(3, 2), (2, -1), (1, 1)]:
        flm = oracle.project_sh(Y[l * l + m + l], Y, 3, th, dth, dph)
        for l2 in range(4):
            for m2 in range(-l2, l2 + 1):
                if m2 != m:
                    assert abs(flm[l2 * l2 + m2 + l2]) < 1e-14, (l, m, l2, m2)


def test_project_sh_edge_weights():
    """A constant integrates to 4pi to the theta trapezoid error; the half weights
    at theta = 0, pi and phi = 0, 2pi are what make the sum converge."""
    th, ph, dth, dph = _sph_grid(201, 9)
    Y = oracle.sph_harmonics(0, th, ph)
    f = np.ones((len(th), len(ph)), dtype=complex)
    flm = oracle.project_sh(f, Y, 0, th, dth, dph)
    # integral of conj(Y00) = 4pi * Y00 = sqrt(4pi), up to the pole rows (sin = 0 there)
    assert abs(flm[0] - math.sqrt(4 * math.pi)) < 1e-4


# ------------------------------------------------------------------ Bessel basis, radial transform
def test_bessel_basis_kind0_is_orthonormal_on_zero_to_rmax():
    R = 10.0
    r = np.linspace(0.0, R, 2001)
    kv, jb = oracle.bessel_basis(3, 6, r, R, kind=0)
    dr = r[1] - r[0]
    w = np.full(len(r), dr)
    w[0] = w[-1] = 0.5 * dr
    for l in range(4):
        for n in range(6):
            assert kv[l, n] == pytest.approx(oracle.jl_zero(l, n + 1) / R, rel=1e-15)
        G = (jb[l] * r ** 2 * w) @ jb[l].T
        assert np.abs(np.diag(G) - 1.0).max() < 1e-13      # normalised by its own trapezoid
        off = G - np.diag(np.diag(G))
        assert np.abs(off).max() < 1e-4, l
        # the normalisation matches the closed form R^3/2 j_{l+1}(z)^2
        z = oracle.jl_zero(l, 2)
        raw = spherical_jn(l, z * r / R)
        want = raw / math.sqrt(R ** 3 / 2 * spherical_jn(l + 1, z) ** 2)
        assert np.abs(jb[l, 1] - want).max() < 1e-5


def test_bessel_basis_kind1_values():
    r = np.linspace(0.5, 20.0, 300)
    kv, jb = oracle.bessel_basis(8, 16, r, 20.0, kind=1, kmax=4.0)
    for l in range(9):
        np.testing.assert_allclose(kv[l], 4.0 * np.arange(1, 17) / 16, rtol=0, atol=0)
        ref = spherical_jn(l, kv[l][:, None] * r[None, :])
        assert np.abs(jb[l] - ref).max() < 1e-12


def test_radial_project_closed_form_and_linearity():
    """f_00(r) = 1: ft(n) = int_0^R j_0(k r) r^2 dr = (sin kR - kR cos kR)/k^3,
    approached at O(dr^2) by the trapezoid."""
    lmax, nk, R = 1, 8, 6.0
    errs = []
    for nr in (301, 601):
        r = np.linspace(0.0, R, nr)
        kv, jb = oracle.bessel_basis(lmax, nk, r, R, kind=1, kmax=3.0)
        flm = np.zeros((nr, (lmax + 1) ** 2), dtype=complex)
        flm[:, 0] = 1.0
        ft = oracle.radial_project(lmax, nk, r, r[1] - r[0], jb, flm)
        k = kv[0]
        want = (np.sin(k * R) - k * R * np.cos(k * R)) / k ** 3
        errs.append(np.abs(ft[:, 0] - want).max())
        assert np.abs(ft[:, 1:]).max() == 0.0
        # linearity in flm
        rng = np.random.default_rng(3)
        a = rng.normal(size=flm.shape) + 1j * rng.normal(size=flm.shape)
        b = rng.normal(size=flm.shape) + 1j * rng.normal(size=flm.shape)
        fa = oracle.radial_project(lmax, nk, r, r[1] - r[0], jb, a)
        fb = oracle.radial_project(lmax, nk, r, r[1] - r[0], jb, b)
        fab = oracle.radial_project(lmax, nk, r, r[1] - r[0], jb, 2.0 * a - b)
        assert np.abs(fab - (2.0 * fa - fb)).max() < 1e-10 * np.abs(fa).max()
    assert 3.5 < errs[0] / errs[1] < 4.5, errs


# ------------------------------------------------------------------ domain_to_sph
def _padded_coords(Nx, Ny, Nz, g, xmin, xmax):
    dx = (xmax - xmin) / Nx
    dy = math.pi / (Ny - 1)
    dz = 2 * math.pi / (Nz - 1)
    x = xmin + np.arange(Nx + 2 * g + 1) * dx
    y = -dy / 2 + (np.arange(Ny + 2 * g + 1) - g) * dy
    z = np.arange(Nz + 1) * dz
    return x, y, z, dx, dy, dz


@pytest.mark.parametrize("shape", [(12, 9, 17), (20, 17, 9)])
def test_domain_to_sph_is_exact_on_trilinear_functions(shape):
    Nx, Ny, Nz = shape
    g, xmin, xmax = 2, 1.8, 9.0
    x, y, z, dx, dy, dz = _padded_coords(Nx, Ny, Nz, g, xmin, xmax)
    Z, Yc, X = np.meshgrid(z, y, x, indexing="ij")
    rs = xmin + np.arange(Nx) * dx
    ts = np.arange(Ny) * dy
    ps = np.arange(Nz) * dz
    P, T, Rr = np.meshgrid(ps, ts, rs, indexing="ij")
    funcs = [
        (lambda a, b, c: a, lambda a, b, c: a),
        (lambda a, b, c: b, lambda a, b, c: b),
        (lambda a, b, c: c, lambda a, b, c: c),
        (lambda a, b, c: 0.3 + 2 * a - b * c + 0.5 * a * b * c, None),
    ]
    for fpad, _ in funcs:
        fd = fpad(X, Yc, Z)
        out = oracle.domain_to_sph(fd, Nx, Ny, Nz, g, xmin, dx, dy, dz)
        want = fpad(Rr, T, P)
        assert np.abs(out - want).max() < 1e-12 * max(1.0, np.abs(want).max())


def test_domain_to_sph_radial_and_phi_nodes_are_copied():
    """Spectral r and phi nodes coincide with evolution nodes (r_i = xmin + i dx is
    padded column i, since the padded x starts at xmin), so the result is the
    theta average of the two bracketing cell-centred rows."""
    Nx, Ny, Nz, g = 10, 9, 13, 2
    x, y, z, dx, dy, dz = _padded_coords(Nx, Ny, Nz, g, 1.8, 7.0)
    rng = np.random.default_rng(11)
    fd = rng.normal(size=(len(z), len(y), len(x)))
    out = oracle.domain_to_sph(fd, Nx, Ny, Nz, g, 1.8, dx, dy, dz)
    for k in range(Nz):
        for j in range(Ny):
            jj = j + g  # padded row with y <= theta_j < next row
            a, b = fd[k, jj, 0:Nx], fd[k, jj + 1, 0:Nx]
            np.testing.assert_allclose(out[k, j], 0.5 * (a + b), rtol=0, atol=1e-14)
