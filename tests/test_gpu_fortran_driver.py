"""The Fortran side of the boundary: a driver shaped like the reference's
evolve loop (main/evolve.f90:22-26) calls the engine through
kg_amd/fortran/kg_hip_mod.f90 (ISO_C_BINDING) on the reference's own state
layout u(5, -g:Nx+g, -g:Ny+g, 0:Nz) complex(8), and its result equals the
oracle's."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from kg_amd.initial_data import shell_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "kg_amd", "fortran", "kg_evolve")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built (no amdflang)")]


@pytest.mark.parametrize("scheme", ["pirk2", "rk4"])
def test_fortran_driver_matches_oracle(tmp_path, scheme):
    Nx, Ny, Nz, xmin, xmax, m = 20, 9, 17, 1.8, 10.0, 0.2
    re = shell_state(Nx, Ny, Nz, 2, xmin, xmax, 5.0, 1.2, 1.0, {(2, 0): 1.0, (1, -1): 0.4})
    im = shell_state(Nx, Ny, Nz, 2, xmin, xmax, 6.0, 1.0, 0.3, {(1, 1): 1.0})
    u = re + 1j * im                                     # [5][k][j][i]
    uf = np.asfortranarray(np.transpose(u, (0, 3, 2, 1)))  # u(5, i, j, k), Fortran order
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    uf.ravel(order="F").tofile(fin)
    nsteps = 6
    print("fortran driver", scheme, flush=True)
    r = subprocess.run([EXE, str(Nx), str(Ny), str(Nz), repr(xmin), repr(xmax), repr(m), scheme, str(nsteps),
                        str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(fout, dtype=np.complex128).reshape(uf.shape, order="F")
    got = np.transpose(got, (0, 3, 2, 1))
    for part, arr in (("re", re), ("im", im)):
        sim = oracle.Sim(Nx, Ny, Nz, 2, xmin, xmax, "Kerr_Schild", 0.0, 1.0, m, "outflow", 0.25)
        sim.u[...] = arr
        sim.step(scheme, nsteps)
        g = got.real if part == "re" else got.imag
        assert np.array_equal(g, sim.u), (scheme, part)
