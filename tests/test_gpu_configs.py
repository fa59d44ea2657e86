"""The workload configurations' physics on the GPU, beyond the small-grid cases.

- The full single-GPU workload's physics (m = 0.1, r in [1.8, 200], a Gaussian
  shell r0 = 30, width 3, times sum_{l<=4} c_lm Y_lm with c_lm ~ N(0, 1e-3),
  RK4 with the 4th-order stencils) on a reduced grid, 200 steps against the
  oracle; and at its full size (2048 x 128 x 256) a few steps whose result is
  the same bit for bit under other launch shapes and in a batched launch.
- The PIRK/RK3 workload's base physics (m = 0.3, an l = 1, m = 1 pulse at
  r0 = 8, width 1.5, on r in [1.8, 100]): PIRK2, PIRK3 and RK3 for 1000 steps
  on a reduced grid against the oracle.
- Members of the batched population's sizes (Nr = 256 with each angular class)
  in one batch, against the oracle.
- Angular self-convergence: an l = 2, m = 1 pulse with r, theta and phi
  refined together converges at the operator set's order away from the poles
  and the phi seam.
"""
import numpy as np
import pytest

import oracle
from kg_amd import LinesMethod
from kg_amd.batch import MemberSpec, batch_step, make_member
from kg_amd.initial_data import random_lm_coeffs, shell_state

pytestmark = pytest.mark.gpu


def _rel_err(a, b):
    out = []
    for c in range(5):
        den = np.abs(b[c]).max()
        out.append(float(np.abs(a[c] - b[c]).max() / den) if den > 0 else float(np.abs(a[c]).max()))
    return out


def test_full_workload_physics_reduced_grid_200_steps():
    """Nr = 512, Ntheta = 32, Nphi = 64 on r in [1.8, 200] with the full
    workload's mass, pulse, operators and Sommerfeld outer boundary:
    bit-identical after 200 RK4 steps."""
    Nx, Ny, Nz = 512, 33, 65
    print("full-workload physics on 512 x 32 x 64, fd4 RK4, 200 steps", flush=True)
    sim = oracle.Sim(Nx, Ny, Nz, 2, 1.8, 200.0, "Kerr_Schild", 0.0, 1.0, 0.1, "sommerfeld", 0.25)
    sim.set_order(4)
    u0 = sim.fill(shell_state(Nx, Ny, Nz, 2, 1.8, 200.0, 30.0, 3.0, 1.0, random_lm_coeffs(20261203)))
    sim.u[...] = u0
    gpu = LinesMethod(Nx=Nx, Ny=Ny, Nz=Nz, xmin=1.8, xmax=200.0, CFL=0.25, m=0.1, scheme="rk4", operators="fd4",
                      boundary_type_rmax="sommerfeld")
    assert gpu.dt == sim.dt
    gpu.set_u_soa(u0[None])
    sim.step("rk4", 200)
    gpu.step(200)
    got = gpu.u_soa()[0]
    gpu.close()
    assert np.isfinite(got).all() and np.abs(got[0]).max() > 1e-4
    assert max(_rel_err(got, sim.u)) <= 1e-12
    assert np.array_equal(got, sim.u)


def test_full_workload_full_size_is_shape_and_batch_invariant():
    """2048 x 128 x 256, fd4 RK4, m = 0.1, Sommerfeld outer boundary: 20 steps
    at the default launch shape,
    at two others and as a one-member batched launch give the same state bit
    for bit, finite and nonzero."""
    import torch

    from kg_amd.initial_data import shell_state_device

    Nx, Ny, Nz = 2048, 129, 257
    args = (Nx, Ny, Nz, 2, 1.8, 200.0, 30.0, 3.0, 1.0, random_lm_coeffs(20261203))
    u0 = shell_state_device(*args)
    torch.cuda.synchronize()
    results = []
    for label, kc in (("default", None), ("kc=8", 8), ("kc=64", 64), ("batched", None)):
        print("full size, 20 steps,", label, flush=True)
        lm = LinesMethod(Nx=Nx, Ny=Ny, Nz=Nz, xmin=1.8, xmax=200.0, CFL=0.25, m=0.1, scheme="rk4",
                         operators="fd4", boundary_type_rmax="sommerfeld")
        lm.copy_state_device(u0.data_ptr(), to_handle=True)
        lm.boundary()
        if kc is not None:
            lm.set_tuning(kc, 8)
        if label == "batched":
            batch_step([lm], 20)
        else:
            lm.step(20)
        out = torch.empty_like(u0)
        lm.sync()
        lm.copy_state_device(out.data_ptr(), to_handle=False)
        torch.cuda.synchronize()
        lm.close()
        results.append(out)
    ref = results[0]
    assert bool(torch.isfinite(ref).all())
    assert float(ref[0].abs().max()) > 1e-4
    assert not torch.equal(ref, u0)
    for r in results[1:]:
        assert torch.equal(r, ref)


@pytest.mark.parametrize("scheme,operators", [("pirk2", "ref2"), ("pirk3", "ref2"), ("rk3", "ref2"),
                                              ("rk3", "fd4")])
def test_pirk_workload_base_physics_1000_steps(scheme, operators):
    """m = 0.3, l = 1, m = 1 Gaussian pulse (r0 = 8, width 1.5) on r in
    [1.8, 100] with Nr = 256, Ntheta = 16, Nphi = 32: 1000 steps bit-identical
    to the oracle, PIRK stage arrays included."""
    Nx, Ny, Nz = 256, 17, 33
    print("PIRK workload base physics:", scheme, operators, "1000 steps", flush=True)
    sim = oracle.Sim(Nx, Ny, Nz, 2, 1.8, 100.0, "Kerr_Schild", 0.0, 1.0, 0.3, "outflow", 0.25)
    sim.set_order(4 if operators == "fd4" else 2)
    u0 = sim.fill(shell_state(Nx, Ny, Nz, 2, 1.8, 100.0, 8.0, 1.5, 1e-3, {(1, 1): 1.0}))
    sim.u[...] = u0
    gpu = LinesMethod(Nx=Nx, Ny=Ny, Nz=Nz, xmin=1.8, xmax=100.0, CFL=0.25, m=0.3, scheme=scheme,
                      operators=operators)
    gpu.set_u_soa(u0[None])
    for _ in range(4):
        sim.step(scheme, 250)
        gpu.step(250)
    got = gpu.u_soa()[0]
    assert np.isfinite(got).all() and np.abs(got[0]).max() > 1e-6
    assert max(_rel_err(got, sim.u)) <= 1e-12
    assert np.array_equal(got, sim.u)
    if scheme in ("pirk2", "pirk3"):
        assert np.array_equal(gpu.u_stage_soa(0)[0], sim.u_stage)
        if scheme == "pirk3":
            assert np.array_equal(gpu.u_stage_soa(1)[0], sim.u_stage_2)
    gpu.close()


@pytest.mark.parametrize("operators", ["ref2", "fd4"])
def test_population_sized_members_in_one_batch(operators):
    """One member of each angular class of the batched population at Nr = 256
    (the population's smallest radial size), with masses and black holes drawn
    as there: 10 batched RK4 steps equal the oracle bit for bit."""
    from kg_amd.batch import DR
    specs = []
    for q, (Nt, Np) in enumerate([(16, 32), (32, 64), (64, 128)]):
        M = [0.7, 1.4, 1.9][q]
        xmin = 0.9 * 2.0 * M
        specs.append(MemberSpec(256, Nt, Np, [0.05, 0.3, 0.45][q], M, xmin, xmin + 256 * DR, xmin + 14.0 + 3 * q,
                                1.5, random_lm_coeffs(40 + q), 40 + q))
    print("population-sized batch,", operators, flush=True)
    batch = [make_member(s, operators=operators) for s in specs]
    u0 = [b.u_soa()[0] for b in batch]
    batch_step(batch, 10)
    for q, s in enumerate(specs):
        sim = oracle.Sim(s.Nr, s.Ntheta + 1, s.Nphi + 1, 2, s.xmin, s.xmax, "Kerr_Schild", 0.0, s.M_bh, s.m,
                         "outflow", 0.25)
        sim.set_order(4 if operators == "fd4" else 2)
        sim.u[...] = u0[q]
        sim.step("rk4", 10)
        got = batch[q].u_soa()[0]
        assert np.abs(got - u0[q]).max() > 0
        assert np.array_equal(got, sim.u), q
        batch[q].close()


def _angular_run(Nr, n_steps, dt, operators):
    Ny, Nz = Nr // 4 + 1, Nr // 2 + 1
    xmin, xmax = 6.0, 26.0
    lm = LinesMethod(Nx=Nr, Ny=Ny, Nz=Nz, xmin=xmin, xmax=xmax, dt=dt, m=0.2, scheme="rk4", operators=operators)
    u0 = shell_state(Nr, Ny, Nz, 2, xmin, xmax, 16.0, 2.0, 1.0, {(2, 1): 1.0})
    lm.set_u_soa(u0[None])
    lm.boundary()
    lm.step(n_steps)
    u = lm.u_soa()[0]
    lm.close()
    return u


def _theta_to_coarse(u, Ny_coarse, q):
    """Values of a grid refined q times (q = 2, 4) in theta at the coarse cell
    centres of its physical rows j = 1 .. Ny_coarse-1: the coarse centre
    (j - 1/2) q h lies midway between fine rows q j - q/2 and q j - q/2 + 1,
    and the 4-point cubic (-1, 9, 9, -1)/16 over rows q j - q/2 - 1 ..
    q j - q/2 + 2 is 4th-order accurate.  u is padded (row j at index j + 2)."""
    g = 2
    out = []
    for j in range(1, Ny_coarse):
        jl = q * j - q // 2 + g
        out.append((-u[..., jl - 1, :] + 9.0 * u[..., jl, :] + 9.0 * u[..., jl + 1, :] - u[..., jl + 2, :]) / 16.0)
    return np.stack(out, axis=-2)


@pytest.mark.parametrize("operators,lo,hi", [("ref2", 3.5, 4.5), ("fd4", 13.0, 19.0)])
def test_angular_self_convergence(operators, lo, hi):
    """l = 2, m = 1 shell (r0 = 16, width 2) on Kerr-Schild M = 1, m = 0.2,
    r in [6, 26]; (Nr, Ntheta, Nphi) = (96, 24, 48), (192, 48, 96),
    (384, 96, 192) with dt in proportion to the smallest cell (the polar row
    at r = 6: dt ~ h^2, so the time error is negligible); T = 2.  Compared on
    r in [10, 22], theta in [0.5, pi - 0.5], phi in [0.5, 2 pi - 0.5] -- out of
    causal reach of the poles and the seam within T."""
    T = 2.0
    n0 = 160
    dt0 = T / n0
    runs = []
    for lvl, Nr in enumerate((96, 192, 384)):
        print("angular convergence", operators, "Nr", Nr, flush=True)
        runs.append(_angular_run(Nr, n0 * 4 ** lvl, dt0 / 4 ** lvl, operators))
    # common points: r index ii (coarse) = 2 ii (mid) = 4 ii (fine); phi index
    # k = 2k = 4k; theta by 4th-order interpolation to the coarse centres
    Ny0 = 96 // 4 + 1
    mid = _theta_to_coarse(runs[1], Ny0, 2)
    fine = _theta_to_coarse(runs[2], Ny0, 4)
    coarse = runs[0][..., 3:2 + Ny0, :]  # physical rows 1 .. Ny0-1
    dx = 20.0 / 96
    ii = np.arange(int(np.ceil((10.0 - 6.0) / dx)), int((22.0 - 6.0) / dx) + 1)
    dth = np.pi / (Ny0 - 1)
    th = (np.arange(1, Ny0) - 0.5) * dth
    jj = np.nonzero((th >= 0.5) & (th <= np.pi - 0.5))[0]
    dph = 2 * np.pi / (96 // 2)
    kk = np.nonzero((np.arange(96 // 2 + 1) * dph >= 0.5) & (np.arange(96 // 2 + 1) * dph <= 2 * np.pi - 0.5))[0]
    sel = np.ix_(range(5), kk, jj, ii)
    a = coarse[sel]
    b = mid[:, 2 * kk][:, :, jj][..., 2 * ii]
    c = fine[:, 4 * kk][:, :, jj][..., 4 * ii]
    for f in range(5):
        e1 = np.abs(a[f] - b[f]).max()
        e2 = np.abs(b[f] - c[f]).max()
        print(operators, "field", f, "differences", e1, e2, "ratio", e1 / e2, flush=True)
        assert lo < e1 / e2 < hi, (operators, f, e1 / e2)
