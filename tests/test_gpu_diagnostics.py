"""On-device diagnostics against the oracle (MoL/diagnostic_lib.f90:185-274).

Point values (density, currents, their interpolation) are formed with the
reference's arithmetic; the integrals and sphere sums add in a different
order than the oracle's nested loops, so they are held to 1e-12 of the
integral of the absolute integrand (the FP64 bar).
"""
import math

import numpy as np
import pytest

import oracle
from kg_amd import LinesMethod
from kg_amd.diagnostics import detector_indices, detectors, integrate
from kg_amd.initial_data import random_lm_coeffs, shell_state
from kg_amd.spectral import SpectralProjector

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _setup(Nx, Ny, Nz, xmin, xmax, a, m, cplx, seed):
    lm = LinesMethod(Nx=Nx, Ny=Ny, Nz=Nz, xmin=xmin, xmax=xmax, a=a, m=m, scheme="rk4", is_complex=cplx)
    sim = oracle.Sim(Nx, Ny, Nz, 2, xmin, xmax, "Kerr_Schild", a, 1.0, m)
    rng = np.random.default_rng(seed)
    parts = []
    for p in range(lm.nparts):
        u = shell_state(Nx, Ny, Nz, 2, xmin, xmax, 0.5 * (xmin + xmax), 0.25 * (xmax - xmin), 1.0,
                        random_lm_coeffs(seed + p, 3, 1.0))
        parts.append(u + 0.05 * rng.normal(size=u.shape))
    lm.set_u_soa(np.stack(parts))
    lm.boundary()
    return lm, sim


CASES = [
    (24, 9, 17, 1.8, 1.8 + 24 * 0.35, 0.0, 0.3, False),
    (40, 17, 33, 1.08, 1.08 + 40 * 0.3, 0.6, 0.1, True),
    (17, 7, 13, 3.42, 3.42 + 17 * 0.4, 0.0, 0.0, False),
]


def _grid(Nx, Ny, Nz, xmin, xmax):
    dx = ((xmin + 1.0 * (xmax - xmin)) - xmin) / Nx
    return dx, math.acos(-1.0) / (Ny - 1), 2.0 * math.acos(-1.0) / (Nz - 1)


@pytest.mark.parametrize("case", CASES)
def test_integrate_fields_and_density(case):
    Nx, Ny, Nz, xmin, xmax, a, m, cplx = case
    lm, sim = _setup(Nx, Ny, Nz, xmin, xmax, a, m, cplx, 3)
    u = lm.u_soa()
    print("integrate", case, flush=True)
    for f in range(5):
        got = integrate(lm, f)
        want = oracle.integrate(sim, u[0, f])
        scale = oracle.integrate(sim, np.abs(u[0, f]))
        if cplx:
            want = want + 1j * oracle.integrate(sim, u[1, f])
            scale += oracle.integrate(sim, np.abs(u[1, f]))
        assert abs(got - want) <= TOL * scale, (f, got, want)
    rho = oracle.density(sim, u[0], u[1] if cplx else None)
    sg = np.broadcast_to(sim.table("sqrtg")[None], rho.shape)
    want = oracle.integrate(sim, rho * sg)
    got = integrate(lm, "density")
    assert abs(got - want) <= TOL * oracle.integrate(sim, np.abs(rho * sg)), (got, want)


@pytest.mark.parametrize("case", CASES)
def test_detectors_match_oracle(case):
    Nx, Ny, Nz, xmin, xmax, a, m, cplx = case
    lm, sim = _setup(Nx, Ny, Nz, xmin, xmax, a, m, cplx, 4)
    lmax = 4
    sp = SpectralProjector(lm, lmax=lmax, nk=4)
    dx, dy, dz = _grid(Nx, Ny, Nz, xmin, xmax)
    r_spec = xmin + np.arange(Nx) * dx
    idx, _ = detector_indices(np.linspace(xmin, xmin + (Nx - 1.5) * dx, 4), r_spec)
    mom, alm = detectors(sp, idx)
    u = lm.u_soa()
    j = oracle.currents(sim, u[0], u[1] if cplx else None)
    sg = sim.table("sqrtg")[None]
    th, ph = np.arange(Ny) * dy, np.arange(Nz) * dz
    Y = oracle.sph_harmonics(lmax, th, ph)
    phi_s = oracle.domain_to_sph(u[0, 0], Nx, Ny, Nz, 2, xmin, dx, dy, dz).astype(complex)
    if cplx:
        phi_s = phi_s + 1j * oracle.domain_to_sph(u[1, 0], Nx, Ny, Nz, 2, xmin, dx, dy, dz)
    print("detectors", case, idx, flush=True)
    for c in range(3):
        js = j[c] * sg
        f = oracle.domain_to_sph(js.real, Nx, Ny, Nz, 2, xmin, dx, dy, dz) + \
            1j * oracle.domain_to_sph(js.imag, Nx, Ny, Nz, 2, xmin, dx, dy, dz)
        for d, i in enumerate(idx):
            want = oracle.trapezium_2d_complex(f[:, :, i].T, dy, dz)
            scale = oracle.trapezium_2d_complex(np.abs(f[:, :, i].T), dy, dz).real
            assert abs(mom[d, c] - want) <= TOL * max(scale, 1e-300), (c, d, mom[d, c], want)
    for d, i in enumerate(idx):
        want = oracle.project_sh(phi_s[:, :, i].T, Y, lmax, th, dy, dz)
        assert np.abs(alm[d] - want).max() <= TOL * np.abs(want).max(), d

