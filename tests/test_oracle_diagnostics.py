"""Pins of the diagnostics oracle (oracle/kg_diag_oracle.c): density(),
currents_ADM, integrate and trapezium_2D_complex of MoL/diagnostic_lib.f90
and tools/integral.f90, against closed forms and exactness properties."""
import math

import numpy as np
import pytest

import oracle


def _sim(m=0.3, a=0.0):
    return oracle.Sim(12, 9, 9, 2, 1.8, 6.0, "Kerr_Schild", a, 1.0, m)


def _close(actual, desired, **kw):
    np.testing.assert_allclose(actual, np.broadcast_to(desired, np.shape(actual)), **kw)


def test_integrate_is_exact_on_trilinear_functions():
    s = _sim()
    x, y, z = s.table("x"), s.table("y"), s.table("z")
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    g = s.g
    x0, x1 = x[g], x[s.ni - 1 - g]
    y0, y1 = y[g], y[s.nj - 1 - g]
    z0, z1 = z[0], z[-1]
    V = (x1 - x0) * (y1 - y0) * (z1 - z0)
    assert oracle.integrate(s, np.ones_like(X)) == pytest.approx(V, rel=1e-14)
    f = 0.5 + 2.0 * X - Y + 3.0 * Z + X * Y * Z
    want = V * (0.5 + (x0 + x1) - 0.5 * (y0 + y1) + 1.5 * (z0 + z1)) + \
        (x1 ** 2 - x0 ** 2) / 2 * (y1 ** 2 - y0 ** 2) / 2 * (z1 ** 2 - z0 ** 2) / 2
    assert oracle.integrate(s, f) == pytest.approx(want, rel=1e-13)


def test_integrate_converges_at_second_order():
    errs = []
    for n in (16, 32, 64):
        s = oracle.Sim(n, 5, 5, 2, 1.0, 3.0, "Kerr_Schild", 0.0, 1.0, 0.0)
        x = s.table("x")
        f = np.broadcast_to(np.sin(x)[None, None, :], (s.nk, s.nj, s.ni)).copy()
        g = s.g
        x0, x1 = x[g], x[s.ni - 1 - g]
        Vyz = (s.table("y")[s.nj - 1 - g] - s.table("y")[g]) * (s.table("z")[-1] - s.table("z")[0])
        errs.append(abs(oracle.integrate(s, f) - Vyz * (math.cos(x0) - math.cos(x1))))
    assert 3.6 < errs[0] / errs[1] < 4.4 and 3.6 < errs[1] / errs[2] < 4.4, errs


def test_density_terms():
    s = _sim(m=0.3)
    grr, grp, gtt, gpp = (s.table(n)[None] for n in ("g_rr", "g_rp", "g_tt", "g_pp"))
    u = np.zeros(s.shape)
    u[4] = -2.0  # |Pi| enters linearly
    assert np.allclose(oracle.density(s, u), 1.0, rtol=0, atol=1e-15)
    u[:] = 0.0
    u[1] = 3.0
    _close(oracle.density(s, u), 0.5 * grr * 9.0, rtol=1e-15)
    u[:] = 0.0
    u[0] = -4.0  # m^2 |phi|
    _close(oracle.density(s, u), 0.5 * 0.09 * 4.0, rtol=1e-14)
    u[:] = 0.0
    u[2], u[3] = 2.0, -1.0
    _close(oracle.density(s, u), 0.5 * (gtt * 4.0 + gpp * 1.0), rtol=1e-14)
    # the cross term 2 g^{r phi} |psi_r| psi_phi (spinning hole: g^{r phi} != 0)
    s2 = _sim(m=0.0, a=0.7)
    grp2 = s2.table("g_rp")[None]
    assert np.abs(grp2).max() > 0
    u = np.zeros(s2.shape)
    u[1], u[3] = -1.0, 2.0
    grr2, gpp2 = s2.table("g_rr")[None], s2.table("g_pp")[None]
    _close(oracle.density(s2, u), 0.5 * (grr2 + 2 * grp2 * 2.0 + 4.0 * gpp2), rtol=1e-13)


def test_density_of_complex_state_uses_moduli_and_real_part():
    s = _sim(m=0.5)
    rng = np.random.default_rng(1)
    ur, ui = rng.normal(size=s.shape), rng.normal(size=s.shape)
    z = ur + 1j * ui
    grr, grp, gtt, gpp = (s.table(n)[None] for n in ("g_rr", "g_rp", "g_tt", "g_pp"))
    want = 0.5 * (np.abs(z[4]) + grr * np.abs(z[1]) ** 2 + (2 * grp * np.abs(z[1]) * z[3]).real
                  + gtt * np.abs(z[2]) ** 2 + gpp * np.abs(z[3]) ** 2 + 0.25 * np.abs(z[0]))
    _close(oracle.density(s, ur, ui), want, rtol=1e-13)


def test_currents_closed_form_schwarzschild():
    """a = 0: g^{r phi} = 0, g^{rr} = 1/(1+2M/r), g^{thth} = 1/r^2."""
    s = _sim()
    r = s.table("x")[None, None, :]
    u = np.zeros(s.shape)
    u[4], u[1], u[2], u[3] = 2.0, 3.0, 5.0, 7.0
    j = oracle.currents(s, u)
    _close(j[0].real, -2.0 * 3.0 / (1 + 2.0 / r) * np.ones((s.nk, s.nj, 1)), rtol=1e-13)
    _close(j[1].real, -2.0 * 5.0 / r ** 2 * np.ones((s.nk, s.nj, 1)), rtol=1e-13)
    gpp = s.table("g_pp")[None]
    _close(j[2].real, -2.0 * 7.0 * gpp * np.ones((s.nk, 1, 1)), rtol=1e-13)
    assert np.abs(j.imag).max() == 0.0


def test_currents_complex_product():
    s = _sim(a=0.5)
    rng = np.random.default_rng(2)
    ur, ui = rng.normal(size=s.shape), rng.normal(size=s.shape)
    z = ur + 1j * ui
    g11, g13, g22, g33 = (s.table(n)[None] for n in ("g_rr", "g_rp", "g_tt", "g_pp"))
    j = oracle.currents(s, ur, ui)
    _close(j[0], -z[4] * (g11 * z[1] + g13 * z[3]), rtol=1e-12, atol=1e-14)
    _close(j[1], -z[4] * g22 * z[2], rtol=1e-12, atol=1e-14)
    _close(j[2], -z[4] * (g13 * z[1] + g33 * z[3]), rtol=1e-12, atol=1e-14)


def test_trapezium_2d_complex_weights():
    f = np.ones((7, 11)) * (1 + 2j)
    assert oracle.trapezium_2d_complex(f, 0.1, 0.2) == pytest.approx((1 + 2j) * 6 * 0.1 * 10 * 0.2, rel=1e-14)
    # exact on bilinear functions
    a, b = np.meshgrid(np.arange(7) * 0.1, np.arange(11) * 0.2, indexing="ij")
    assert oracle.trapezium_2d_complex(a * b, 0.1, 0.2).real == pytest.approx((0.6 ** 2 / 2) * (2.0 ** 2 / 2),
                                                                              rel=1e-13)
