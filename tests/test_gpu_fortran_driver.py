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


@pytest.mark.parametrize("scheme,batch,bc", [("pirk2", 1, "outflow"), ("rk4", 1, "outflow"), ("rk4_fd4", 1, "outflow"),
                                             ("pirk2", 3, "outflow"), ("rk4_fd4", 2, "outflow"),
                                             ("pirk3", 1, "sommerfeld"), ("rk4_fd4", 2, "sommerfeld")])
def test_fortran_driver_matches_oracle(tmp_path, scheme, batch, bc):
    """batch > 1: the driver advances that many copies through
    kg_lines_method_batch_step (kg_batch_step) and checks they agree; bc: the
    outer boundary the driver passes to the constructor."""
    Nx, Ny, Nz, xmin, xmax, m = 20, 9, 17, 1.8, 10.0, 0.2
    re = shell_state(Nx, Ny, Nz, 2, xmin, xmax, 5.0, 1.2, 1.0, {(2, 0): 1.0, (1, -1): 0.4})
    im = shell_state(Nx, Ny, Nz, 2, xmin, xmax, 6.0, 1.0, 0.3, {(1, 1): 1.0})
    u = re + 1j * im                                     # [5][k][j][i]
    uf = np.asfortranarray(np.transpose(u, (0, 3, 2, 1)))  # u(5, i, j, k), Fortran order
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    uf.ravel(order="F").tofile(fin)
    nsteps = 6
    print("fortran driver", scheme, "batch", batch, flush=True)
    extra = (["batch=%d" % batch] if batch > 1 else []) + (["bc=" + bc] if bc != "outflow" else [])
    r = subprocess.run([EXE, str(Nx), str(Ny), str(Nz), repr(xmin), repr(xmax), repr(m), scheme, str(nsteps),
                        str(fin), str(fout)] + extra, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    if batch > 1:
        assert "batch of %4d members agree" % batch in r.stdout, r.stdout
    got = np.fromfile(fout, dtype=np.complex128).reshape(uf.shape, order="F")
    got = np.transpose(got, (0, 3, 2, 1))
    for part, arr in (("re", re), ("im", im)):
        sim = oracle.Sim(Nx, Ny, Nz, 2, xmin, xmax, "Kerr_Schild", 0.0, 1.0, m, bc, 0.25)
        if scheme.endswith("_fd4"):
            sim.set_order(4)
        sim.u[...] = arr
        sim.step(scheme.replace("_fd4", ""), nsteps)
        g = got.real if part == "re" else got.imag
        assert np.array_equal(g, sim.u), (scheme, part)


@pytest.mark.parametrize("kind,nsteps", [("random_shell", 0), ("spherical_shell", 4)])
def test_fortran_driver_initial_data(tmp_path, kind, nsteps):
    """MoL%initial_conditions through the binding (random phases from seed 77),
    then nsteps RK4 steps, against the oracle (the random spectrum's exp and
    cos/sin round differently on the device: agreement to 1e-12 of each field)."""
    Nx, Ny, Nz, xmin, xmax, m = 24, 9, 9, 1.8, 8.0, 0.5
    mu, sigma, ncart, seed = 4.5, 0.9, 32, 77
    fout = tmp_path / "out.bin"
    print("fortran driver initial data", kind, nsteps, flush=True)
    r = subprocess.run([EXE, str(Nx), str(Ny), str(Nz), repr(xmin), repr(xmax), repr(m), "rk4", str(nsteps),
                        "ic=%s,%r,%r,%d,%d" % (kind, mu, sigma, ncart, seed), str(fout)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "initial data " + kind in r.stdout
    shape = (5, Nx + 5, Ny + 5, Nz + 1)
    got = np.transpose(np.fromfile(fout, dtype=np.complex128).reshape(shape, order="F"), (0, 3, 2, 1))
    sim = oracle.Sim(Nx, Ny, Nz, 2, xmin, xmax, "Kerr_Schild", 0.0, 1.0, m, "outflow", 0.25)
    r_max = sim.table("x")[-1]
    field = oracle.random_field(ncart, r_max, 1.0, oracle.uniforms(seed, ncart)) if kind != "spherical_shell" else None
    want, amp = oracle.initial_conditions(sim, kind, mu, sigma, field, r_max)
    line = [ln for ln in r.stdout.splitlines() if ln.startswith("initial data")][0]
    assert float(line.split("Amp")[1]) == pytest.approx(amp, rel=1e-13)
    parts = []
    for arr in (want.real, want.imag):
        sim.u[...] = arr
        sim.step("rk4", nsteps)
        parts.append(sim.u.copy())
    want = parts[0] + 1j * parts[1]
    for f in range(5):
        scale = max(np.abs(want[f]).max(), 1e-300)
        assert np.abs(got[f] - want[f]).max() <= 1e-12 * scale, (kind, f)


@pytest.mark.parametrize("scheme,bc", [("pirk2", "outflow"), ("rk3_fd4", "sommerfeld")])
def test_fortran_driver_refined_matches_oracle(tmp_path, scheme, bc):
    """The two-level refinement through the Fortran binding
    (kg_refined_lines_method): base and fine levels read from and written to
    the reference's u(5, i, j, k) layout, both equal to the oracle's
    refinement (oracle/refinement.py) bit for bit after 8 base steps."""
    from oracle.refinement import RefinedSim

    Nx, Ny, Nz, xmin, xmax, m, r_patch = 40, 9, 13, 1.8, 12.0, 0.2, 5.0
    order = 4 if scheme.endswith("_fd4") else 2
    sims = [RefinedSim(Nx, Ny, Nz, 2, xmin, xmax, "Kerr_Schild", 0.0, 1.0, m, bc, 0.25, r_patch, order=order)
            for _ in range(2)]
    files = {}
    for lv_name in ("base", "fine"):
        parts = []
        for q, R in enumerate(sims):
            lv = getattr(R, lv_name)
            u0 = shell_state(lv.Nx, lv.Ny, lv.Nz, 2, xmin, xmin + lv.Nx * lv.dx, 4.5 + q, 1.2, 1.0 - 0.6 * q,
                             {(1, 1): 1.0, (2, 0): 0.3})
            if order == 4:
                u0 = lv.fill(u0.copy())
            lv.u[...] = u0
            parts.append(u0)
        u = parts[0] + 1j * parts[1]
        uf = np.asfortranarray(np.transpose(u, (0, 3, 2, 1)))
        fin, fout = tmp_path / (lv_name + "_in.bin"), tmp_path / (lv_name + "_out.bin")
        uf.ravel(order="F").tofile(fin)
        files[lv_name] = (fin, fout, uf.shape)
    nsteps = 8
    print("fortran driver refined", scheme, bc, flush=True)
    extra = ["refine=%r,%s,%s" % (r_patch, files["fine"][0], files["fine"][1])] + (["bc=" + bc] if bc != "outflow"
                                                                                    else [])
    r = subprocess.run([EXE, str(Nx), str(Ny), str(Nz), repr(xmin), repr(xmax), repr(m), scheme, str(nsteps),
                        str(files["base"][0]), str(files["base"][1])] + extra, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "refined steps" in r.stdout, r.stdout
    for R in sims:
        R.step(scheme.replace("_fd4", ""), nsteps)
    for lv_name in ("base", "fine"):
        fin, fout, shape = files[lv_name]
        got = np.fromfile(fout, dtype=np.complex128).reshape(shape, order="F")
        got = np.transpose(got, (0, 3, 2, 1))
        assert np.array_equal(got.real, getattr(sims[0], lv_name).u), (lv_name, "re")
        assert np.array_equal(got.imag, getattr(sims[1], lv_name).u), (lv_name, "im")
