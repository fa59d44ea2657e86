"""GPU parity of the initial data (main/initial_conditions.f90:23-188,
tools/random.f90:17-156) against the oracle.

The spherical shell, the FFT butterflies, the cell search and the stencils
repeat the reference's operations exactly; the random spectrum's exp and
cos/sin come from the device library, so the random types agree to rounding
(|GPU - oracle| <= 1e-12 of each field's largest value), as does Amp, whose
integral sums in a different order (1e-13 relative).
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from kg_amd import KGError, LinesMethod
from kg_amd import _capi as capi
from kg_amd.initial_data import initial_conditions, random_field

pytestmark = pytest.mark.gpu

GRID = dict(Nx=24, Ny=9, Nz=9, g=2, xmin=1.8, xmax=8.0, m=0.5)


def _pair(is_complex, bc="outflow", scheme="rk4", metric="Kerr_Schild", a=0.0, **kw):
    p = dict(GRID)
    p.update(kw)
    lm = LinesMethod(g_pts=p["g"], xmin=p["xmin"], xmax=p["xmax"], Nx=p["Nx"], Ny=p["Ny"], Nz=p["Nz"], m=p["m"],
                     boundary_type_rmax=bc, scheme=scheme, is_complex=is_complex, CFL=0.25, metric=metric, a=a)
    s = oracle.Sim(p["Nx"], p["Ny"], p["Nz"], p["g"], p["xmin"], p["xmax"], metric, a, 1.0, p["m"], bc)
    return lm, s


def _state(lm):
    u = lm.u_soa()
    return u[0] + 1j * u[1] if lm.is_complex else u[0].astype(np.complex128)


def _assert_fields_close(got, want, rel=1e-12):
    for f in range(5):
        scale = max(np.abs(want[f]).max(), 1e-300)
        err = np.abs(got[f] - want[f]).max()
        assert err <= rel * scale, (f, err, scale)


@pytest.mark.parametrize("n,r_max,p0", [(16, 5.0, 1.0), (32, 9.0, 0.7), (64, 20.0, 1.0)])
def test_random_field_matches_oracle(n, r_max, p0):
    print("random field", n, r_max, p0, flush=True)
    seed = 1000 + n
    U = oracle.uniforms(seed, n)
    got = random_field(n, r_max, p0, seed=seed)
    assert np.array_equal(got, random_field(n, r_max, p0, uniforms=U))
    want = oracle.random_field(n, r_max, p0, U)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()


def test_random_field_at_the_reference_size():
    """ncart = 256, the reference's fixed Cartesian grid, against numpy's FFT."""
    n, r_max = 256, 61.0
    print("random field 256", flush=True)
    got = random_field(n, r_max, 1.0, seed=5)
    U = oracle.uniforms(5, n)
    R = r_max + 1.0
    w = 2 * np.pi * np.fft.fftfreq(n, d=2 * R / n)
    psi_p = np.exp(-0.5 * (w[None, None, :] ** 2 + w[None, :, None] ** 2 + w[:, None, None] ** 2)) \
        * np.exp(2j * np.pi * U)
    del U
    want = np.fft.fftn(psi_p) / n ** 3
    del psi_p
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()


@pytest.mark.parametrize("kind,is_complex,bc", [
    ("spherical_shell", False, "outflow"), ("spherical_shell", True, "outflow"),
    ("random", True, "outflow"), ("random", True, "inflow"),
    ("random_shell", True, "outflow"), ("random_shell", True, "inflow")])
def test_initial_data_matches_oracle(kind, is_complex, bc):
    print("initial data", kind, is_complex, bc, flush=True)
    lm, s = _pair(is_complex, bc)
    n, seed, mu, sigma = 32, 77, 4.5, 0.9
    r_max = s.table("x")[-1]
    amp = initial_conditions(lm, kind, mu, sigma, ncart=n, seed=seed)
    field = oracle.random_field(n, r_max, 1.0, oracle.uniforms(seed, n)) if kind != "spherical_shell" else None
    want, amp_o = oracle.initial_conditions(s, kind, mu, sigma, field, r_max)
    assert amp == pytest.approx(amp_o, rel=1e-13)
    got = _state(lm)
    _assert_fields_close(got, want)
    if not is_complex:
        assert np.all(want.imag == 0.0)
    lm.close()


@pytest.mark.parametrize("kind,metric,a", [("spherical_shell", "Schw_EF", 0.0), ("random_shell", "Kerr_Schild", 0.7)])
def test_initial_data_other_backgrounds(kind, metric, a):
    """The lapse, shift and sqrt(gamma) of the handle's metric enter Pi and the
    normalising integrals (Schwarzschild in EF form, spinning Kerr-Schild)."""
    print("initial data", kind, metric, a, flush=True)
    xmin = 2.5 if metric == "Schw_EF" else 1.8
    lm, s = _pair(True, metric=metric, a=a, xmin=xmin)
    n, seed, mu, sigma = 32, 5, 5.0, 1.0
    r_max = s.table("x")[-1]
    amp = initial_conditions(lm, kind, mu, sigma, ncart=n, seed=seed)
    field = oracle.random_field(n, r_max, 1.0, oracle.uniforms(seed, n)) if kind != "spherical_shell" else None
    want, amp_o = oracle.initial_conditions(s, kind, mu, sigma, field, r_max)
    assert amp == pytest.approx(amp_o, rel=1e-13)
    _assert_fields_close(_state(lm), want)
    lm.close()


def test_reference_workload_random_shell():
    """The small reference workload's grid (Nx=256, Ny=33, Nz=65, r to 60)
    with the reference's ncart = 256."""
    print("random_shell on the 256x33x65 grid", flush=True)
    lm, s = _pair(True, Nx=256, Ny=33, Nz=65, xmin=1.8, xmax=60.0, m=0.5)
    n, seed = 256, 3
    r_max = s.table("x")[-1]
    amp = initial_conditions(lm, "random_shell", 15.0, 2.0, ncart=n, seed=seed)
    field = oracle.random_field(n, r_max, 1.0, oracle.uniforms(seed, n))
    want, amp_o = oracle.initial_conditions(s, "random_shell", 15.0, 2.0, field, r_max)
    del field
    assert amp == pytest.approx(amp_o, rel=1e-13)
    _assert_fields_close(_state(lm), want)
    lm.close()


def test_initial_data_then_evolve():
    """random_shell initial data, then 20 RK4 steps, on both sides."""
    print("random_shell + 20 rk4 steps", flush=True)
    lm, s = _pair(True, scheme="rk4")
    n, seed = 32, 11
    r_max = s.table("x")[-1]
    initial_conditions(lm, "random_shell", 4.5, 1.0, ncart=n, seed=seed)
    field = oracle.random_field(n, r_max, 1.0, oracle.uniforms(seed, n))
    want, _ = oracle.initial_conditions(s, "random_shell", 4.5, 1.0, field, r_max)
    # evolve the oracle's real and imaginary systems
    parts = []
    for part in (want.real, want.imag):
        so = oracle.Sim(GRID["Nx"], GRID["Ny"], GRID["Nz"], GRID["g"], GRID["xmin"], GRID["xmax"], "Kerr_Schild",
                        0.0, 1.0, GRID["m"], "outflow", 0.25, lm.dt)
        so.u[...] = part
        so.step("rk4", 20)
        parts.append(so.u.copy())
        so.close()
    lm.step(20)
    _assert_fields_close(_state(lm), parts[0] + 1j * parts[1], rel=1e-11)
    lm.close()


def test_rejections():
    lm, _ = _pair(False)
    with pytest.raises(KGError) as ei:
        initial_conditions(lm, "random", ncart=32)
    assert ei.value.status == capi.KG_ERR_INVALID and "complex" in str(ei.value)
    lm.close()
    lm, _ = _pair(True)
    with pytest.raises(KGError) as ei:  # r_max + 1 = 11.2 >= 16/2
        initial_conditions(lm, "random_shell", 4.0, 1.0, ncart=16)
    assert ei.value.status == capi.KG_ERR_INVALID and "ncart" in str(ei.value)
    with pytest.raises(KGError):
        initial_conditions(lm, "MPS_test")
    with pytest.raises(KGError) as ei:
        initial_conditions(lm, "spherical_shell", 4.0, 0.0)
    assert ei.value.status == capi.KG_ERR_INVALID
    lib = capi.load()
    d = capi.KGICDesc(type=7, ncart=32, mu=1.0, sigma=1.0, p0=1.0)
    assert lib.kg_initial_data(lm._h, C.byref(d), None) == capi.KG_ERR_INVALID
    lm.close()
