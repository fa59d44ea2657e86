"""Self-convergence of the GPU path at three resolutions.

A spherically symmetric massive pulse on Schwarzschild in Kerr-Schild
coordinates stays spherical (every angular difference of an angle-independent
field vanishes), so only the radial stencils (2nd order,
finite_differences.f90:17-44,767-801) and the time integrator (RK4, 4th order)
carry error.  With dt proportional to dx the solution converges at 2nd order:
the ratio of successive differences |u_N - u_2N| / |u_2N - u_4N| at the common
radii approaches 2^2 = 4.  A 4th-order radial operator set would give 16.
"""
import numpy as np
import pytest

from kg_amd import LinesMethod
from kg_amd.initial_data import shell_state

pytestmark = pytest.mark.gpu


def _run(Nx, T, xmin=1.8, xmax=41.8):
    Ny, Nz = 5, 9
    # dt = CFL * min(dx, e^xmin dtheta, ...) (diagnostic_lib.f90:101-104): dx is the minimum here
    lm = LinesMethod(Nx=Nx, Ny=Ny, Nz=Nz, xmin=xmin, xmax=xmax, CFL=0.25, m=0.2, scheme="rk4")
    dt = lm.dt
    assert abs(dt - 0.25 * (xmax - xmin) / Nx) < 1e-12
    u0 = shell_state(Nx, Ny, Nz, 2, xmin, xmax, 15.0, 2.0, 1.0, {(0, 0): 1.0})
    lm.set_u_soa(u0[None])
    lm.boundary()
    n = int(round(T / dt))
    assert abs(n * dt - T) < 1e-9
    lm.step(n)
    u = lm.u_soa()[0]
    lm.close()
    return u


def test_rk4_ref2_converges_at_second_order():
    T = 6.0
    coarse = [_run(N, T) for N in (160, 320, 640)]
    # padded column ii sits at x = xmin + ii dx, so coarse column ii is column 2 ii
    # of the next grid and 4 ii of the finest; compare on r in [5, 35], away from
    # the excision and outer boundaries
    dx = 40.0 / 160
    cols = np.arange(int(np.ceil((5.0 - 1.8) / dx)), int((35.0 - 1.8) / dx) + 1)
    a = coarse[0][:, :, :, cols]
    b = coarse[1][:, :, :, 2 * cols]
    c = coarse[2][:, :, :, 4 * cols]
    for f in (0, 1, 4):  # phi, psi_r, Pi (psi_theta, psi_phi stay zero)
        e1 = np.abs(a[f] - b[f]).max()
        e2 = np.abs(b[f] - c[f]).max()
        q = e1 / e2
        print("field", f, "differences", e1, e2, "ratio", q, flush=True)
        assert 3.5 < q < 4.5, (f, q)
    assert np.abs(coarse[2][2]).max() < 1e-12 and np.abs(coarse[2][3]).max() < 1e-12
