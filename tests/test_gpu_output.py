"""Snapshot arrays formed on the GPU (kg_diag_output) against the state read
back and the oracle's density(), and a snapshot written in the reference's
HDF5 layout (main/save_gnuplot.f90:121-144)."""
import os

import numpy as np
import pytest

import oracle
from kg_amd import LinesMethod
from kg_amd.initial_data import random_lm_coeffs, shell_state
from kg_amd.output import HDF5, SNAPSHOT, output_array, snapshot_coordinates, write_snapshot

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cplx", [False, True])
def test_output_arrays(cplx):
    Nx, Ny, Nz = 24, 9, 17
    lm = LinesMethod(Nx=Nx, Ny=Ny, Nz=Nz, xmin=1.8, xmax=10.2, a=0.5, m=0.3, scheme="rk4", is_complex=cplx)
    u0 = shell_state(Nx, Ny, Nz, 2, 1.8, 10.2, 6.0, 1.5, 1.0, random_lm_coeffs(3, 3, 0.5))
    parts = [u0, -0.7 * u0[::-1]] if cplx else [u0]
    lm.set_u_soa(np.stack(parts))
    u = lm.u_soa()
    z = u[0] + (1j * u[1] if cplx else 0)
    g = 2
    sl = (slice(0, Nz), slice(g, g + Ny + 1), slice(0, Nx + g + 1))
    want = {0: np.abs(z[0]), 2: np.abs(z[1]), 3: np.abs(z[2]), 4: np.abs(z[3]), 5: np.abs(z[4]), 6: z[0].real,
            7: z[0].imag if cplx else np.zeros_like(z[0].real)}
    for q, w in want.items():
        got = output_array(lm, q)
        np.testing.assert_allclose(got, w[sl], rtol=5e-16, atol=0, err_msg=str(q))
    sim = oracle.Sim(Nx, Ny, Nz, 2, 1.8, 10.2, "Kerr_Schild", 0.5, 1.0, 0.3)
    rho = oracle.density(sim, u[0], u[1] if cplx else None)
    np.testing.assert_allclose(output_array(lm, 1), rho[sl], rtol=4e-15, atol=0)


def test_snapshot_files(tmp_path):
    try:
        h5 = HDF5()
    except RuntimeError as e:
        pytest.skip(str(e))
    Nx, Ny, Nz = 20, 9, 13
    lm = LinesMethod(Nx=Nx, Ny=Ny, Nz=Nz, xmin=1.8, xmax=9.0, m=0.1, scheme="rk4")
    lm.set_u_soa(shell_state(Nx, Ny, Nz, 2, 1.8, 9.0, 5.0, 1.5, 1.0, random_lm_coeffs(4, 2, 0.5))[None])
    write_snapshot(lm, 0, str(tmp_path), h5)
    lm.step(3)
    write_snapshot(lm, 1, str(tmp_path), h5)
    X, Y, Z = snapshot_coordinates(lm)
    for q, name in SNAPSHOT:
        fn = os.path.join(str(tmp_path), name + ".h5")
        assert np.array_equal(h5.read(fn, "/refinement_1/Xcoord"), X)
        assert np.array_equal(h5.read(fn, "/refinement_1/Zcoord"), Z)
        assert np.array_equal(h5.read(fn, "/refinement_1/%s_1" % name), output_array(lm, q))
        assert h5.exists(fn, "/refinement_1/%s_0" % name)
