"""Two-level refinement on the GPU (kg_refined_*) against the oracle's
refinement (oracle/refinement.py): both levels bit-identical after many base
steps, for every scheme and operator set, with the PIRK/RK3 workload's physics
(m = 0.3, an l = 1, m = 1 pulse at r0 = 8, width 1.5, r in [1.8, 100], a 2:1
patch on [1.8, 10]) on a reduced grid for 1000 base steps."""
import numpy as np
import pytest

from kg_amd.initial_data import random_lm_coeffs, shell_state
from kg_amd.refinement import RefinedLinesMethod
from oracle.refinement import RefinedSim

pytestmark = pytest.mark.gpu


def _pair(Nr, Ny, Nz, xmin, xmax, r_patch, scheme, operators, m=0.3, a=0.0, M=1.0, bc="outflow", coeffs=None,
          r0=8.0, w=1.5, cplx=False, seed=0):
    order = 4 if operators == "fd4" else 2
    osims = [RefinedSim(Nr, Ny, Nz, 2, xmin, xmax, "Kerr_Schild", a, M, m, bc, 0.25, r_patch, order=order)
             for _ in range(2 if cplx else 1)]
    gpu = RefinedLinesMethod(r_patch, Nx=Nr, Ny=Ny, Nz=Nz, xmin=xmin, xmax=xmax, CFL=0.25, m=m, a=a, M_bh=M,
                             boundary_type_rmax=bc, scheme=scheme, operators=operators, is_complex=cplx)
    rng = np.random.default_rng(seed)
    for lv_name in ("base", "fine"):
        parts = []
        for q, R in enumerate(osims):
            lv = getattr(R, lv_name)
            u0 = shell_state(lv.Nx, lv.Ny, lv.Nz, 2, xmin, xmin + lv.Nx * lv.dx, r0, w, 1e-3 * (1 + q),
                             coeffs or {(1, 1): 1.0})
            u0[4] += 1e-5 * rng.standard_normal(u0[4].shape)
            if order == 4:
                u0 = lv.fill(u0.copy())
            lv.u[...] = u0
            parts.append(u0)
        getattr(gpu, lv_name).set_u_soa(np.stack(parts))
    return osims, gpu


def _check(osims, gpu, scheme):
    gb, gf = gpu.base.u_soa(), gpu.fine.u_soa()
    for q, R in enumerate(osims):
        assert np.isfinite(gb[q]).all() and np.abs(gb[q][0]).max() > 0
        assert np.array_equal(gb[q], R.base.u), ("base", q)
        assert np.array_equal(gf[q], R.fine.u), ("fine", q)
    if scheme in ("pirk2", "pirk3") and len(osims) == 1:
        assert np.array_equal(gpu.base.u_stage_soa(0)[0], osims[0].base.u_stage)
        assert np.array_equal(gpu.fine.u_stage_soa(0)[0], osims[0].fine.u_stage)


SCHEMES = [("pirk2", "ref2"), ("pirk3", "ref2"), ("rk3", "ref2"), ("rk4", "ref2"), ("rk3", "fd4"), ("rk4", "fd4")]


@pytest.mark.parametrize("scheme,operators", SCHEMES)
def test_refined_steps_match_oracle(scheme, operators):
    """A small refined pair around a spinning hole with a Sommerfeld base
    boundary and ragged data: 12 base steps (24 fine), in two calls."""
    print("refined", scheme, operators, flush=True)
    osims, gpu = _pair(48, 9, 13, 1.9, 14.0, 6.0, scheme, operators, a=0.6, bc="sommerfeld",
                       coeffs=random_lm_coeffs(3, 2, 0.3), r0=5.0, seed=1)
    for R in osims:
        R.step(scheme, 12)
    gpu.step(5)
    gpu.step(7)
    _check(osims, gpu, scheme)
    assert gpu.t == pytest.approx(12 * osims[0].base.dt, rel=1e-14)
    gpu.close()


@pytest.mark.parametrize("scheme,operators", [("pirk2", "ref2"), ("rk4", "fd4")])
def test_refined_complex_state(scheme, operators):
    osims, gpu = _pair(40, 7, 9, 1.8, 12.0, 5.0, scheme, operators, cplx=True, r0=4.5)
    for R in osims:
        R.step(scheme, 6)
    gpu.step(6)
    _check(osims, gpu, scheme)
    gpu.close()


@pytest.mark.parametrize("scheme,operators", [("pirk2", "ref2"), ("pirk3", "ref2"), ("rk3", "ref2"), ("rk3", "fd4")])
def test_pirk_workload_refined_1000_base_steps(scheme, operators):
    """The PIRK/RK3 workload's physics with its 2:1 patch on [1.8, 10] on a
    reduced base grid (Nr = 256 on [1.8, 100], Ntheta = 16, Nphi = 32):
    1000 base steps (2000 fine) bit-identical to the oracle."""
    print("refined PIRK workload", scheme, operators, flush=True)
    osims, gpu = _pair(256, 17, 33, 1.8, 100.0, 10.0, scheme, operators)
    for _ in range(4):
        for R in osims:
            R.step(scheme, 250)
        gpu.step(250)
        print("  250 base steps more", flush=True)
    _check(osims, gpu, scheme)
    gpu.close()


@pytest.mark.parametrize("scheme", ["pirk2", "pirk3", "rk3"])
def test_pirk_workload_full_size_refined_is_shape_invariant(scheme):
    """The PIRK/RK3 workload at its full size: base 1024 x 64 x 128 on
    r in [1.8, 100] (m = 0.3, the l = 1, m = 1 pulse at r0 = 8, width 1.5)
    with its 2:1 patch on [1.8, 10]: 100 base steps (200 fine) under two
    launch shapes give the same two levels bit for bit, finite and nonzero."""
    import torch

    from kg_amd.initial_data import shell_state_device

    Nx, Ny, Nz, xmin, xmax = 1024, 65, 129, 1.8, 100.0
    dx0 = ((xmin + 1.0 * (xmax - xmin)) - xmin) / Nx
    results = []
    for kc, tj in ((8, 8), (32, 4)):
        print("full-size refined", scheme, "kc", kc, "tj", tj, flush=True)
        gpu = RefinedLinesMethod(10.0, Nx=Nx, Ny=Ny, Nz=Nz, xmin=xmin, xmax=xmax, CFL=0.25, m=0.3,
                                 boundary_type_rmax="outflow", scheme=scheme)
        for lv in (gpu.base, gpu.fine):
            lv_xmax = xmin + lv.Nx * (dx0 if lv is gpu.base else dx0 / 2.0)
            u0 = shell_state_device(lv.Nx, lv.Ny, lv.Nz, 2, xmin, lv_xmax, 8.0, 1.5, 1e-3, {(1, 1): 1.0})
            torch.cuda.synchronize()
            lv.copy_state_device(u0.data_ptr(), to_handle=True)
            lv.boundary()
            lv.set_tuning(kc, tj)
        gpu.step(40)
        gpu.step(60)
        out = []
        for lv in (gpu.base, gpu.fine):
            buf = torch.empty((5, lv.nk, lv.nj, lv.ni), dtype=torch.float64, device="cuda")
            lv.sync()
            lv.copy_state_device(buf.data_ptr(), to_handle=False)
            torch.cuda.synchronize()
            out.append(buf)
        gpu.close()
        results.append(out)
    for lvl in range(2):
        a, b = results[0][lvl], results[1][lvl]
        assert bool(torch.isfinite(a).all())
        assert float(a[0].abs().max()) > 1e-6
        assert torch.equal(a, b), ("level", lvl)
