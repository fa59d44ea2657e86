"""Pins of the initial-data oracle (oracle/kg_ic_oracle.c): the Cartesian
random field of tools/random.f90:17-59 against numpy's FFT, its trilinear
sampling (random.f90:72-156) against exactness on trilinear polynomials, and
the three initial_conditions types (main/initial_conditions.f90:23-188)
against closed forms, the normalisation each one imposes and the relations
between the fields."""
import math

import numpy as np
import pytest

import oracle


def _trapz3(f, s):
    """integrate (integral.f90:646-669) by numpy: trapezoid over ii in [g, ni-1-g],
    jj in [g, nj-1-g] and every phi plane."""
    g = s.g
    x, y, z = s.table("x"), s.table("y"), s.table("z")
    sub = f[:, g:s.nj - g, g:s.ni - g]
    a = np.trapezoid(sub, x=x[g:s.ni - g], axis=2)
    a = np.trapezoid(a, x=y[g:s.nj - g], axis=1)
    return np.trapezoid(a, x=z, axis=0)


def _omega(n, r_max):
    R = r_max + 1.0
    dx = 2 * R / n
    return 2 * np.pi * np.fft.fftfreq(n, d=dx)


@pytest.mark.parametrize("n,r_max,p0", [(8, 2.5, 1.0), (16, 5.0, 1.0), (32, 9.0, 0.7)])
def test_random_field_is_numpy_fftn_over_n3(n, r_max, p0):
    U = oracle.uniforms(7 + n, n)
    got = oracle.random_field(n, r_max, p0, U)
    w = _omega(n, r_max)
    W2 = w[None, None, :] ** 2 + w[None, :, None] ** 2 + w[:, None, None] ** 2
    psi_p = np.exp(-0.5 / p0 ** 2 * W2) * np.exp(2j * np.pi * U)
    want = np.fft.fftn(psi_p) / n ** 3
    assert np.abs(got - want).max() <= 1e-14 * np.abs(want).max()
    # a plausible slip -- the inverse with the other sign -- is far away
    assert np.abs(got - np.fft.ifftn(psi_p)).max() > 1e-3 * np.abs(want).max()


def test_uniforms_are_splitmix64():
    # SplitMix64 (Steele, Lea, Flood 2014) from state 0: the first outputs of
    # the published generator, top 53 bits
    known = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    got = oracle.uniforms(0, 2)  # 8 values from seed 0
    for k, v in enumerate(known):
        assert got.reshape(-1)[k] == (v >> 11) * 2.0 ** -53
    u = oracle.uniforms(12345, 16).reshape(-1)
    assert u.min() >= 0.0 and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.02


def _sim_for(n, **kw):
    args = dict(Nx=24, Ny=9, Nz=9, g=2, xmin=1.8, xmax=8.0, metric="Kerr_Schild", a=0.0, M_bh=1.0, m=0.5)
    args.update(kw)
    s = oracle.Sim(args["Nx"], args["Ny"], args["Nz"], args["g"], args["xmin"], args["xmax"], args["metric"],
                   args["a"], args["M_bh"], args["m"], args.get("bc", "outflow"))
    r_max = s.table("x")[-1]
    assert r_max + 1.0 < n / 2
    return s, r_max


def _cart(s):
    x, y, z = s.table("x"), s.table("y"), s.table("z")
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X * np.sin(Y) * np.cos(Z), X * np.sin(Y) * np.sin(Z), X * np.cos(Y)


def test_sampling_is_exact_on_trilinear_polynomials():
    n = 32
    s, r_max = _sim_for(n)
    R = r_max + 1.0
    nodes = -R + np.arange(n) * (2 * R / n)

    def P(x, y, z):
        return (1.0 + 2.0 * x - 3.0 * y + 0.5 * z + 0.25 * x * y - 0.125 * y * z + 0.0625 * x * z
                + 0.03 * x * y * z)

    Zn, Yn, Xn = np.meshgrid(nodes, nodes, nodes, indexing="ij")
    field = P(Xn, Yn, Zn) + 1j * P(-Yn, Zn, Xn)
    got = oracle.sample(s, field, r_max)
    xc, yc, zc = _cart(s)
    want = P(xc, yc, zc) + 1j * P(-yc, zc, xc)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    # one cell off (a plausible search mistake) breaks it
    field2 = np.roll(field, 1, axis=2)
    assert np.abs(oracle.sample(s, field2, r_max) - want).max() > 1e-3


def test_spherical_shell_closed_form():
    s, _ = _sim_for(64)
    mu, sigma = 4.5, 0.8
    u, amp = oracle.initial_conditions(s, "spherical_shell", mu, sigma)
    assert np.all(u.imag == 0.0)
    u = u.real
    x = s.table("x")
    env = np.exp(-((x - mu) ** 2) / sigma ** 2)
    phi = env / x
    dtphi = -2.0 * (x - mu) / (x * sigma ** 2) * env
    alpha, sqrtg = s.table("alpha"), s.table("sqrtg")
    np.testing.assert_allclose(u[0], amp * np.broadcast_to(phi, u[0].shape), rtol=1e-15, atol=0)
    assert np.all(u[1] == 0.0) and np.all(u[3] == 0.0)
    # psi_theta: differences of equal values vanish, except the one-sided last
    # row, where 3v - 4v + v leaves the rounding of 3v
    assert np.all(u[2][:, :-1] == 0.0)
    assert np.abs(u[2][:, -1]).max() <= 4 * np.finfo(float).eps * np.abs(u[0]).max() / (2 * s.dy)
    np.testing.assert_allclose(u[4], amp * np.broadcast_to(dtphi[None, :] / alpha, u[4].shape), rtol=1e-14,
                               atol=1e-300)
    # Amp from the unscaled fields: density = 0.5 (|Pi| + m^2 |phi|) (psi = 0)
    rho = 0.5 * (np.abs(dtphi[None, :] / alpha) + s_m(s) ** 2 * np.abs(phi)[None, :])
    I = _trapz3(np.broadcast_to(rho * alpha * sqrtg, (s.nk, s.nj, s.ni)), s)
    assert amp == pytest.approx(1.0 / math.sqrt(I), rel=1e-13)


def s_m(s):
    return 0.5  # the mass _sim_for gives every oracle Sim here


def _d1(f, h, axis):
    """central differences with the reference's one-sided ends (x) -- numpy restatement"""
    d = np.empty_like(f)
    f = np.moveaxis(f, axis, -1)
    dv = np.moveaxis(d, axis, -1)
    dv[..., 1:-1] = (f[..., 2:] - f[..., :-2]) / (2 * h)
    dv[..., 0] = (-3 * f[..., 0] + 4 * f[..., 1] - f[..., 2]) / (2 * h)
    dv[..., -1] = (3 * f[..., -1] - 4 * f[..., -2] + f[..., -3]) / (2 * h)
    return d


@pytest.mark.parametrize("bc", ["outflow", "inflow"])
def test_random_fields_relations_and_norm(bc):
    n = 32
    s, r_max = _sim_for(n, bc=bc)
    U = oracle.uniforms(99, n)
    field = oracle.random_field(n, r_max, 1.0, U)
    u, amp = oracle.initial_conditions(s, "random", 0.0, 1.0, field, r_max)
    m = s_m(s)
    sqrtg, alpha, beta = s.table("sqrtg"), s.table("alpha"), s.table("beta_r")
    # normalisation: integrate(m |phi|^2 sqrt(gamma)) = 1
    assert _trapz3(m * np.abs(u[0]) ** 2 * sqrtg, s) == pytest.approx(1.0, rel=1e-13)
    # interior phi = Amp * sampled field; the x ghost columns are the xL/xR fills
    samp = oracle.sample(s, field, r_max)
    np.testing.assert_allclose(u[0][:, :, 1:-1], amp * samp[:, :, 1:-1], rtol=1e-14, atol=1e-14 * abs(amp))
    want0 = 3 * u[0][:, :, 1] - 3 * u[0][:, :, 2] + u[0][:, :, 3]
    np.testing.assert_allclose(u[0][:, :, 0], want0, rtol=1e-12, atol=1e-13 * np.abs(u[0]).max())
    if bc == "outflow":
        assert np.array_equal(u[0][:, :, -1], u[0][:, :, -2])
    else:
        assert np.all(u[0][:, :, -1] == 0.0)
    # psi_r is the radial stencil derivative of phi (after the x fills)
    np.testing.assert_allclose(u[1], _d1(u[0], s.dx, 2), rtol=0, atol=1e-12 * np.abs(u[1]).max())
    np.testing.assert_allclose(u[3][1:-1], (u[0][2:] - u[0][:-2]) / (2 * s.dz), rtol=0,
                               atol=1e-12 * np.abs(u[3]).max())
    # Pi = 1/alpha (0 - beta psi_r)
    np.testing.assert_allclose(u[4], -(beta / alpha) * u[1], rtol=1e-14, atol=1e-300)


def test_random_shell_fills_and_norm():
    n = 32
    s, r_max = _sim_for(n)
    mu, sigma = 5.0, 1.2
    U = oracle.uniforms(2024, n)
    field = oracle.random_field(n, r_max, 1.0, U)
    u, amp = oracle.initial_conditions(s, "random_shell", mu, sigma, field, r_max)
    m, g = s_m(s), s.g
    sqrtg, alpha, beta = s.table("sqrtg"), s.table("alpha"), s.table("beta_r")
    assert _trapz3(m * np.abs(u[0]) ** 2 * sqrtg, s) == pytest.approx(1.0, rel=1e-13)
    x = s.table("x")
    env = np.exp(-((x - mu) ** 2) / sigma ** 2) / x
    samp = oracle.sample(s, field, r_max)
    inner = (slice(1, s.nk - 1), slice(g + 1, s.nj - g - 1), slice(1, s.ni - 1))
    np.testing.assert_allclose(u[0][inner], (amp * samp * env)[inner], rtol=1e-13, atol=1e-14 * abs(amp))
    # Pi = 1/alpha (psi_r - beta psi_r) away from the ghost rows and columns
    np.testing.assert_allclose(u[4][inner], ((1 - beta) / alpha * u[1])[inner], rtol=1e-13,
                               atol=1e-14 * np.abs(u[4]).max())
    # fills: z periodic copies, then the pole rows taken from phi + pi
    assert np.array_equal(u[:, 0], u[:, s.Nz - 1]) and np.array_equal(u[:, s.Nz], u[:, 1])
    for k in range(s.nk):
        kt = (k + s.Nz // 2) % (s.Nz - 1)
        for gg in range(g, -1, -1):
            assert np.array_equal(u[:, k, g - gg], u[:, kt, gg + 1 + g])
            assert np.array_equal(u[:, k, s.Ny + gg + g], u[:, kt, s.Ny - gg - 1 + g])
    assert np.array_equal(u[:, :, :, -1], u[:, :, :, -2])
