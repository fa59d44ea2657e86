"""Benchmark: Klein-Gordon grid-point updates/s on one MI355X.

Workload (the full single-GPU configuration): massive (m = 0.1) Klein-Gordon
field on Schwarzschild in Kerr-Schild coordinates (M = 1), Nr = 2048 on
r in [1.8, 200], Ntheta = 128, Nphi = 256 (67.1 M points x 5 fields), classic
RK4 with the 4th-order stencil set of tools/finite_differences.f90 (x_4, y_4,
z_4, advec_x_4), dt from CFL 0.25, Sommerfeld outer boundary (--bc outflow:
the reference's copy).  Initial phi is a Gaussian shell (r0 = 30,
width 3) times sum_{l<=4} c_lm Y_lm with c_lm ~ N(0, 1e-3) from a seed; psi =
its gradient, Pi = 0.  The same workload with the reference's live 2nd-order
stencils is timed beside it (`plain_ref2`; --operators ref2 swaps the two).

value = Nr*Ntheta*Nphi * steps / wall time of the timed steps (state resident
in HBM), which replay the step's captured HIP graph as kg_step does in
production.  A separate pass of the same steps, launched eagerly with a HIP
event pair around every kernel on the engine's stream, gives the average
duration of the RHS+stage kernel for the roofline line.

Beside it, on the same GPU: the uneven batched workload (256 evolutions, one
launch per RK4 stage, vs the same grids one at a time) with its
spherical-Bessel projection (l <= 8, 512 k, split BF16 and FP64), the serving
loop (Poisson requests, continuous batching, p50/p99 time-to-result), and
the reference's random_shell initial data built on the GPU for the full grid,
the reference's own production step (PIRK2 with its 2nd-order stencils) on the
full grid, the PIRK/RK3 workload with its 2:1 radial patch, and the CPU oracle
on a bounded sample.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--operators fd4|ref2] [--other-steps K]
                [--pirk2-steps K] [--refined-steps K] [--batched N] [--serve N] [--serve-small N] [--ic N]
                [--ic-only]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
B_PER_UPDATE = {"rk4": 680, "rk3": 320, "pirk2": 248, "pirk3": 416}
# launch shapes (phi chunk, theta rows) the plain path's autotuner chooses from
# for a grid run by itself (with the fd4 set a grid of 256 x 16 x 32 is fastest
# at 2 planes, one of 1024 x 16 x 32 and larger at 8; the fd4 stages always run
# 8 theta rows)
PLAIN_SHAPES = [(kc, tj) for tj in (4, 8) for kc in (1, 2, 4, 8, 16)]
STAGES = {"rk4": 4, "rk3": 3, "pirk2": 3, "pirk3": 5}

CFG = dict(Nx=2048, Ny=129, Nz=257, xmin=1.8, xmax=200.0, m=0.1, M_bh=1.0, a=0.0, CFL=0.25, scheme="rk4",
           bc="sommerfeld")


def cpu_baseline(operators, max_seconds=12.0):
    """The oracle (C restatement, OpenMP over the host threads) on a bounded
    sample of the workload: the same radial/polar grid (Nr = 2048, Ntheta =
    128) with 16 azimuthal planes, RK4 steps with the same operator set until
    ~max_seconds of wall time."""
    import numpy as np
    import oracle
    from kg_amd.initial_data import shell_state, random_lm_coeffs

    Nx, Ny, Nz = CFG["Nx"], CFG["Ny"], 17
    sim = oracle.Sim(Nx, Ny, Nz, 2, CFG["xmin"], CFG["xmax"], "Kerr_Schild", 0.0, 1.0, CFG["m"], CFG["bc"],
                     CFG["CFL"])
    sim.set_order(4 if operators == "fd4" else 2)
    sim.u[...] = shell_state(Nx, Ny, Nz, 2, CFG["xmin"], CFG["xmax"], 30.0, 3.0, 1.0, random_lm_coeffs(20261203))
    steps, t0 = 0, time.perf_counter()
    while True:
        sim.step("rk4", 1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= max_seconds or steps >= 1024:
            break
    ups = Nx * (Ny - 1) * (Nz - 1) * steps / el
    nt = oracle.threads()
    full = CFG["Nx"] * (CFG["Ny"] - 1) * (CFG["Nz"] - 1)
    return {"value": ups, "unit": "grid-point updates/s", "cores": nt, "kind": "port",
            "sample": "oracle/kg_oracle.c RK4 (%s operators, %s outer boundary), %d steps on Nr=%d, Ntheta=%d, "
                      "Nphi=%d (the workload's r-theta grid, 16 phi planes), %d OpenMP threads, %.1f s; the work per "
                      "point update is the same on every phi plane, so at this rate one step of the full %d-point "
                      "grid takes %.2f s" % (operators, CFG["bc"], steps, Nx, Ny - 1, Nz - 1, nt, el, full,
                                             full / ups)}


TRAFFIC_PROFILES = {"ref2": "stage_kernel_rk4_nr2048_traffic.json", "fd4": "stage4_kernel_rk4_nr2048_traffic.json"}


def measured_traffic(tuning, operators):
    """Bytes per stage launch between L2 and the memory fabric, from the
    committed PMC profile of this workload (the TCC read and write request
    counters by request size, exact on streaming kernels of known size); None
    when the launch shape differs from the profiled one."""
    path = os.path.join(ROOT, "profiles", TRAFFIC_PROFILES[operators])
    if not os.path.exists(path):
        return None
    with open(path) as f:
        prof = json.load(f)
    if (prof["tuning"]["kc"], prof["tuning"]["tj"]) != tuple(tuning):
        return None
    return prof["traffic_bytes_per_launch"]


def small_specs(n, seed=20261204):
    """The smallest class of the uneven population alone: Nr = 256, (Ntheta,
    Nphi) = (16, 32), masses and pulses drawn as for the full population."""
    import numpy as np
    from kg_amd.batch import DR, MemberSpec
    from kg_amd.initial_data import random_lm_coeffs

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        m, M = float(rng.uniform(0.0, 0.5)), float(rng.uniform(0.5, 2.0))
        xmin = 0.9 * 2.0 * M
        xmax = xmin + 256 * DR
        r0 = float(rng.uniform(xmin + 10.0, xmin + 0.5 * (xmax - xmin)))
        sd = int(rng.integers(0, 2**31))
        out.append(MemberSpec(256, 16, 32, m, M, xmin, xmax, r0, 1.5, random_lm_coeffs(sd), sd))
    return out


def batched_vs_plain(n, steps, tuning, operators, specs=None, spectral=True):
    """The uneven batched workload: n evolutions (Nr in [256, 2048], (Ntheta, Nphi)
    in {(16,32), (32,64), (64,128)}, m ~ U[0,0.5], M ~ U[0.5,2], RK4), advanced
    `steps` steps (a) all together, one launch per stage (kg_batch_step), and
    (b) one at a time through the plain path (kg_step graph replay per grid,
    each grid at the launch shape its own autotuner picks among PLAIN_SHAPES).
    Aggregate grid-point updates/s of each; inputs resident in HBM."""
    import torch
    from kg_amd.autotune import autotune, autotune_batch
    from kg_amd.batch import batch_step, make_member, uneven_specs

    specs = specs if specs is not None else uneven_specs(n)
    members = []
    for s in specs:
        m = make_member(s, on_device=True, operators=operators)
        members.append(m)
    torch.cuda.empty_cache()
    pts = sum(s.points for s in specs)
    batch_kc, _ = autotune_batch(members, tj=tuning[1])  # includes the warm-up
    t0 = time.perf_counter()
    batch_step(members, steps)
    members[0].sync()
    tb = time.perf_counter() - t0
    # the same batch as persistent launches (wg workgroups per CU march all
    # tiles of a stage in turn), after a warm-up step that builds its plan
    persistent = {}
    for wg in (1, 2):
        members[0].set_batch_persistent(wg)
        batch_step(members, 1)
        members[0].sync()
        t0 = time.perf_counter()
        batch_step(members, steps)
        members[0].sync()
        persistent[str(wg)] = pts * steps / (time.perf_counter() - t0)
    members[0].set_batch_persistent(0)
    # the plain path at its own best shape for each grid (its autotuner, outside
    # the timed region), then a warm-up step (graph capture)
    picks = {}
    for m in members:
        best, _ = autotune(m, candidates=PLAIN_SHAPES, steps=2)
        picks["%d/%d" % best] = picks.get("%d/%d" % best, 0) + 1
        m.step(1)
        m.sync()
    tp = 0.0
    for m in members:
        t0 = time.perf_counter()
        m.step(steps)
        m.sync()
        tp += time.perf_counter() - t0
    spec_out = spectral_projection(members, tb / steps) if spectral else None
    for m in members:
        m.close()
    vb, vp = pts * steps / tb, pts * steps / tp
    return {"members": n, "steps": steps, "operators": operators, "batch_kc": batch_kc,
            "plain_tuning": "autotuned per grid", "plain_picks_kc_tj": dict(sorted(picks.items())),
            "grid_points": pts, "batched_updates_per_s": vb,
            "batched_persistent_updates_per_s": persistent,
            "plain_updates_per_s": vp, "speedup": vb / vp,
            "workload": "uneven RK4 batch: Nr~U{256..2048}, (Ntheta,Nphi) in {(16,32),(32,64),(64,128)}, "
                        "m~U[0,0.5], M~U[0.5,2], seed 20261203",
            "spectral": spec_out}


def spectral_projection(members, batched_step_s):
    """The every-50-steps projection of the batched workload: phi of every member
    onto j_l(k r), l <= 8, k_n = 4n/512 (n = 1..512), split-BF16 matrix cores
    and FP64, results left in HBM: all members in one launch per pass
    (kg_spec_radial_batch), and the same members one projection call each
    (kg_spec_radial on every member's own stream), whose results it must
    equal bit for bit."""
    import numpy as np
    import torch
    from kg_amd.spectral import SpectralProjector, radial_batch

    lmax, nk = 8, 512
    projs = [SpectralProjector(m, lmax=lmax, nk=nk, basis="uniform", kmax=4.0) for m in members]
    nlm = (lmax + 1) ** 2
    outs = {(p, mode): torch.empty((len(projs), nlm * nk * 2), dtype=torch.float64, device="cuda")
            for p in ("split_bf16", "fp64") for mode in ("batch", "each")}
    times = {}
    for rep in range(2):  # the first pass of each warms up
        for prec in ("split_bf16", "fp64"):
            for mode in ("batch", "each"):
                o = outs[(prec, mode)]
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                if mode == "batch":
                    radial_batch(projs, 0, prec, out_device_ptrs=[o[q].data_ptr() for q in range(len(projs))])
                    members[0].sync()
                else:
                    for q, sp in enumerate(projs):
                        sp.radial(0, prec, out_device_ptr=o[q].data_ptr())
                    for m in members:
                        m.sync()
                times[(prec, mode)] = time.perf_counter() - t0
    same = all(torch.equal(outs[(p, "batch")], outs[(p, "each")]) for p in ("split_bf16", "fp64"))
    a, b = outs[("split_bf16", "batch")].cpu().numpy(), outs[("fp64", "batch")].cpu().numpy()
    err = max(float(np.abs(a[q] - b[q]).max() / max(np.abs(b[q]).max(), 1e-300)) for q in range(len(projs)))
    for sp in projs:
        sp.close()
    return {"lmax": lmax, "nk": nk, "members": len(projs),
            "ms_split_bf16": times[("split_bf16", "batch")] * 1e3, "ms_fp64": times[("fp64", "batch")] * 1e3,
            "ms_split_bf16_one_call_per_member": times[("split_bf16", "each")] * 1e3,
            "ms_fp64_one_call_per_member": times[("fp64", "each")] * 1e3,
            "batched_equals_per_member": same,
            "max_rel_err_split_bf16_vs_fp64": err,
            "fraction_of_50_steps": times[("split_bf16", "batch")] / (50.0 * batched_step_s)}


def initial_data(ncart, reps=3):
    """The reference's 'random_shell' initial data (initial_conditions.f90:146-188)
    built on the GPU for the full grid as a complex state: ncart^3 Cartesian random
    field, its FFT, trilinear sampling, envelope, stencils, fills, normalisation."""
    import kg_amd
    from kg_amd.initial_data import initial_conditions

    lm = kg_amd.LinesMethod(Nx=CFG["Nx"], Ny=CFG["Ny"], Nz=CFG["Nz"], xmin=CFG["xmin"], xmax=CFG["xmax"],
                            CFL=CFG["CFL"], m=CFG["m"], M_bh=CFG["M_bh"], a=CFG["a"], scheme=CFG["scheme"],
                            is_complex=True)
    ts = []
    for rep in range(reps + 1):
        t0 = time.perf_counter()
        amp = initial_conditions(lm, "random_shell", 30.0, 3.0, ncart=ncart, seed=rep)
        ts.append(time.perf_counter() - t0)
    lm.close()
    ts = sorted(ts[1:])
    return {"kind": "random_shell", "ncart": ncart, "cartesian_bytes": 16 * ncart ** 3, "ms": ts[len(ts) // 2] * 1e3,
            "amp": amp, "state": "complex, 5 fields on the full grid"}


def serving(nreq, tuning, operators, batch_kc=None, specs=None, workload=None):
    """Serving loop: nreq requests of the batched workload (or of `specs`)
    arriving as a Poisson stream (50/s), steps ~U{500..5000}, admitted against
    the free HBM and run by continuous batching; each returns its final fields
    and its spectrum.  The same requests one at a time through the plain path
    give the comparison."""
    import torch
    from kg_amd.autotune import autotune
    from kg_amd.serving import GpuEngine, ServingLoop, percentile, poisson_requests

    reqs = poisson_requests(nreq, specs=specs)
    eng = GpuEngine(tuning=tuning, operators=operators, batch_kc=batch_kc)
    free, _ = torch.cuda.mem_get_info()
    budget = int(0.85 * free)
    results, stats = ServingLoop(eng, budget, chunk=50).run(reqs, keep_fields=True)
    lat = [r.latency for r in results]
    del results
    # each request alone at the plain path's best shape for its grid (tuned
    # outside the timed region)
    alone = {}
    for r in reqs:
        member = eng.admit(r)
        alone[r.rid], _ = autotune(member[0], candidates=PLAIN_SHAPES, steps=2)
        eng.release(member)
    t0 = time.perf_counter()
    for r in reqs:
        eng.run_alone(r, keep_fields=True, tuning=alone[r.rid])
    tp = time.perf_counter() - t0
    ups_plain = stats.updates / tp
    return {"requests": nreq, "operators": operators, "batch_kc": batch_kc,
            "kc_chunks": {str(k): v for k, v in sorted(eng.kc_chunks.items())},
            "alone_tuning": {str(r.rid): list(alone[r.rid]) for r in reqs},
            "arrival_rate_per_s": 50.0, "hbm_budget_bytes": budget,
            "p50_time_to_result_s": percentile(lat, 50), "p99_time_to_result_s": percentile(lat, 99),
            "served_updates_per_s": stats.updates_per_s, "one_at_a_time_updates_per_s": ups_plain,
            "speedup": stats.updates_per_s / ups_plain, "max_active": stats.max_active,
            "chunks": stats.chunks, "joined_running": stats.joined_running,
            "workload": workload or ("Poisson(50/s) requests drawn as the batched workload, steps ~U{500..5000}, "
                                     "chunks of 50 RK4 steps, final fields + l<=8, 512-k spectrum (split-BF16)")}


PROBE_KINDS = {0: ("read", 1, 0), 1: ("write", 0, 1), 2: ("copy", 1, 1), 3: ("3 reads + 1 write", 3, 1),
               4: ("copy, non-temporal", 1, 1), 5: ("3 reads + 1 write, non-temporal", 3, 1)}


def hbm_probe(n=192 * 2 ** 20, reps=10, blocks=(1024, 2048, 4096, 8192, 16384)):
    """What plain streaming kernels move on this GPU (kg_hbm_probe): FP64
    arrays of n doubles (1.5 GiB each), 16-byte accesses, grid-stride loops,
    `reps` timed launches per launch shape; TB/s = bytes moved / average
    launch time."""
    import ctypes as C
    from kg_amd import _capi as capi

    lib = capi.load()
    out = {}
    for kind, (name, nr, nw) in PROBE_KINDS.items():
        per = {}
        for nb in blocks:
            ms = C.c_double()
            st = lib.kg_hbm_probe(0, kind, n, nb, reps, C.byref(ms))
            if st != 0:
                raise RuntimeError(lib.kg_probe_last_error().decode())
            per[nb] = n * 8 * (nr + nw) / (ms.value * 1e-3) / 1e12
            print("bench: probe %s, %d workgroups: %.3f TB/s" % (name, nb, per[nb]), file=sys.stderr, flush=True)
        best = max(per, key=per.get)
        out[name] = {"best_TB_per_s": per[best], "best_workgroups": best, "frac_of_peak": per[best] * 1e3 / HBM_PEAK_GBS,
                     "TB_per_s_by_workgroups": {str(k): v for k, v in per.items()}}
    return {"array_doubles": n, "reps": reps, "threads_per_workgroup": 256, "kinds": out}


KERNELS = {"ref2": "stage_kernel (RHS + RK4 stage combine, the reference's 2nd-order stencils)",
           "fd4": "stage4_kernel (RHS + RK4 stage combine, the 4th-order stencil set)"}


def refined_workload(scheme, steps, warmup=5):
    """The PIRK/RK3 workload: base 1024 x 64 x 128 on r in [1.8, 100], m = 0.3,
    the l = 1, m = 1 pulse at r0 = 8 (width 1.5), the reference's 2nd-order
    stencils, with its 2:1 patch on [1.8, 10] (two fine steps per base step,
    buffer refill and injection), and the same base grid alone.  Updates
    count every level: base points per base step plus fine points per fine
    step."""
    import torch
    import kg_amd
    from kg_amd.initial_data import shell_state_device
    from kg_amd.refinement import RefinedLinesMethod

    Nx, Ny, Nz, xmin, xmax = 1024, 65, 129, 1.8, 100.0
    dx0 = ((xmin + 1.0 * (xmax - xmin)) - xmin) / Nx
    kw = dict(Nx=Nx, Ny=Ny, Nz=Nz, xmin=xmin, xmax=xmax, CFL=0.25, m=0.3, boundary_type_rmax="outflow",
              scheme=scheme)

    def init(lv, lxmax):
        u0 = shell_state_device(lv.Nx, lv.Ny, lv.Nz, 2, xmin, lxmax, 8.0, 1.5, 1e-3, {(1, 1): 1.0})
        torch.cuda.synchronize()
        lv.copy_state_device(u0.data_ptr(), to_handle=True)
        lv.boundary()
        lv.set_tuning(8, 8)

    print("bench: refined %s workload" % scheme, file=sys.stderr, flush=True)
    gpu = RefinedLinesMethod(10.0, **kw)
    init(gpu.base, xmin + gpu.base.Nx * dx0)
    init(gpu.fine, xmin + gpu.fine.Nx * dx0 / 2.0)
    gpu.step(warmup)
    gpu.base.sync()
    gpu.fine.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gpu.step(steps)
    gpu.base.sync()
    gpu.fine.sync()
    torch.cuda.synchronize()
    t_ref = time.perf_counter() - t0
    base_pts = Nx * (Ny - 1) * (Nz - 1)
    fine_pts = (gpu.fine.Nx - 4) * (Ny - 1) * (Nz - 1)  # physical fine columns, buffer included
    fine_nx = gpu.fine.Nx
    gpu.close()
    alone = kg_amd.LinesMethod(**kw)
    init(alone, xmax)
    alone.step(warmup)
    alone.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    alone.step(steps)
    alone.sync()
    torch.cuda.synchronize()
    t_alone = time.perf_counter() - t0
    alone.close()
    return {"scheme": scheme, "steps": steps, "base_grid": [Nx, Ny - 1, Nz - 1], "fine_nr": fine_nx,
            "refined_ms_per_base_step": t_ref / steps * 1e3,
            "refined_updates_per_s": (base_pts + 2 * fine_pts) * steps / t_ref,
            "unrefined_ms_per_step": t_alone / steps * 1e3,
            "unrefined_updates_per_s": base_pts * steps / t_alone,
            "workload": "base 1024 x 64 x 128, r in [1.8, 100], m=0.3, l=1 m=1 pulse r0=8 w=1.5, 2:1 patch on "
                        "[1.8, 10], reference's 2nd-order stencils, outflow"}


def plain(operators, steps, warmup, kc=0, tj=8, scheme=None):
    """The plain single-grid workload with one operator set (and the workload's
    RK4 unless `scheme` names another integrator): state resident in HBM,
    launch shape autotuned (or fixed), `warmup` untimed steps, `steps` timed
    steps with HIP events around every stage kernel."""
    scheme = scheme or CFG["scheme"]
    import numpy as np
    import torch
    import kg_amd
    from kg_amd.initial_data import random_lm_coeffs, shell_state_device

    print("bench: building initial data (%s)" % operators, file=sys.stderr, flush=True)
    lm = kg_amd.LinesMethod(Nx=CFG["Nx"], Ny=CFG["Ny"], Nz=CFG["Nz"], xmin=CFG["xmin"], xmax=CFG["xmax"],
                            CFL=CFG["CFL"], m=CFG["m"], M_bh=CFG["M_bh"], a=CFG["a"], scheme=scheme,
                            operators=operators, boundary_type_rmax=CFG["bc"])
    u0 = shell_state_device(CFG["Nx"], CFG["Ny"], CFG["Nz"], 2, CFG["xmin"], CFG["xmax"], 30.0, 3.0, 1.0,
                            random_lm_coeffs(20261203))
    torch.cuda.synchronize()
    lm.copy_state_device(u0.data_ptr(), to_handle=True)
    del u0
    torch.cuda.empty_cache()
    lm.boundary()  # ghost-fill the initial data (the reference's IC fills x, y, z ghosts the same way)
    if kc > 0:
        best = (kc, tj)
        lm.set_tuning(*best)
    else:
        from kg_amd.autotune import autotune
        best, timings = autotune(lm, steps=2)
        print("bench: autotune %s %s" % (operators, {"%d/%d" % k: round(v, 4) for k, v in timings.items()}),
              file=sys.stderr, flush=True)
    print("bench: warmup %d steps" % warmup, file=sys.stderr, flush=True)
    lm.step(warmup)
    lm.sync()
    torch.cuda.synchronize()
    # timed region: kg_step as production runs it (one captured HIP graph per
    # step, replayed)
    t0 = time.perf_counter()
    lm.step(steps)
    lm.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # per-kernel durations from a separate pass of the same steps launched
    # eagerly with a HIP event pair around every kernel on the engine's stream
    prof_steps = min(steps, 50)
    lm.profile(True)
    t_p0 = time.perf_counter()
    lm.step(prof_steps)
    lm.sync()
    t_p1 = time.perf_counter()
    stage_ms, stage_n = lm.profile_read(0)
    fill_ms, fill_n = lm.profile_read(1)
    col_ms, col_n = lm.profile_read(2)
    lm.profile(False)
    # the drop-in boundary hands the state over in the reference's layout
    # u(5, i, j, k): one download and one upload, timed (PCIe, host buffers)
    t2 = time.perf_counter()
    uf = lm.u_fortran()
    t3 = time.perf_counter()
    lm.set_u_fortran(uf)
    lm.sync()
    t4 = time.perf_counter()
    finite = bool(np.isfinite(uf).all())
    host_bytes = uf.nbytes
    del uf
    lm.close()
    npts = CFG["Nx"] * (CFG["Ny"] - 1) * (CFG["Nz"] - 1)
    elapsed = t1 - t0
    sch = scheme
    per_launch_bytes = npts * B_PER_UPDATE[sch] / STAGES[sch]
    avg_stage_s = stage_ms / 1e3 / max(stage_n, 1)
    achieved = per_launch_bytes / avg_stage_s / 1e9
    return {"operators": operators, "scheme": sch, "value": npts * steps / elapsed, "ms_per_step": elapsed / steps * 1e3,
            "steps": steps, "state_finite": finite, "tuning": {"kc": best[0], "tj": best[1]},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS,
                         "traffic": measured_traffic(best, operators) if sch == CFG["scheme"] else None,
                         "kernel": KERNELS[operators] if sch == CFG["scheme"] else
                         "stage_kernel (%s stages, the reference's 2nd-order stencils)" % sch.upper(),
                         "avg_kernel_ms": avg_stage_s * 1e3, "launches": stage_n,
                         "bytes_per_launch": per_launch_bytes},
            "kernel_timing": {"steps": prof_steps, "eager_ms_per_step": (t_p1 - t_p0) / prof_steps * 1e3,
                              "stage_ms_per_step": stage_ms / prof_steps, "note":
                              "kernel durations from a separate eager pass with events around each kernel; "
                              "value and ms_per_step time the graph-replay path"},
            "fill_ms_per_step": fill_ms / prof_steps,
            "outer_column_ms_per_step": col_ms / prof_steps,
            "host_transfer": {"bytes": host_bytes, "download_s": t3 - t2, "upload_s": t4 - t3,
                              "pcie_inclusive_updates_per_s": npts * steps / (elapsed + (t3 - t2) + (t4 - t3)),
                              "note": "state handed over once each way in the reference's u(5,i,j,k) layout "
                                      "around the timed steps; not part of value"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--operators", choices=["fd4", "ref2"], default="fd4",
                    help="operator set of the timed workload: fd4 (the 4th-order stencils the workload names) or "
                         "ref2 (the reference's live 2nd-order stencils); the other is timed beside it")
    ap.add_argument("--other-steps", type=int, default=100,
                    help="timed steps of the other operator set (0: skip)")
    ap.add_argument("--pirk2-steps", type=int, default=50,
                    help="timed steps of the reference's PIRK2 with its 2nd-order stencils on the full grid (0: skip)")
    ap.add_argument("--pirk2-kc", type=int, default=0,
                    help="fix the PIRK2 leg's phi planes per workgroup (8 theta rows; skips its autotune)")
    ap.add_argument("--kc", type=int, default=0, help="fix the stage kernel's phi planes per workgroup (skips autotune)")
    ap.add_argument("--tj", type=int, default=8, help="theta rows per workgroup when --kc is given")
    ap.add_argument("--batched", type=int, default=256,
                    help="members of the uneven batched workload to time beside the plain path (0: skip)")
    ap.add_argument("--batched-steps", type=int, default=10)
    ap.add_argument("--refined-steps", type=int, default=200,
                    help="base steps of the PIRK/RK3 workload with its 2:1 patch (PIRK2, PIRK3, RK3); 0 skips it")
    ap.add_argument("--ic", type=int, default=512,
                    help="Cartesian points per axis of the random_shell initial data to time (0: skip)")
    ap.add_argument("--ic-only", action="store_true", help="time only the initial data and exit")
    ap.add_argument("--bc", choices=["sommerfeld", "outflow"], default="sommerfeld",
                    help="outer boundary of the plain single-grid legs (the workload names Sommerfeld)")
    ap.add_argument("--hbm-probe", action="store_true",
                    help="measure what plain streaming kernels move (read, write, copy, 3 reads + 1 write) and exit")
    ap.add_argument("--serve-small", type=int, default=64,
                    help="requests of the population's smallest class through the serving loop (0: skip)")
    ap.add_argument("--serve", type=int, default=32,
                    help="requests of the serving loop to run beside the plain path (0: skip)")
    args = ap.parse_args()
    CFG["bc"] = args.bc
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (a single grid lives on one MI355X); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    assert torch.cuda.is_available(), "no GPU visible"
    if args.ic_only:
        print(json.dumps({"initial_data": initial_data(args.ic or 512)}))
        return
    if args.hbm_probe:
        print(json.dumps({"hbm_probe": hbm_probe()}))
        return
    head = plain(args.operators, args.steps, args.warmup, args.kc, args.tj)
    order = "4th-order" if args.operators == "fd4" else "the reference's 2nd-order"
    out = {
        "metric": "KG grid-point updates/s (batched vs plain, <=1e-12 vs FP64 CPU) + p99 latency",
        "value": head["value"],
        "unit": "grid-point updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Gaussian shell x sum of Y_lm, l<=4)",
        "config": {"workload": "plain single-grid RK4 with %s stencils, Nr=2048 r in [1.8,200], Ntheta=128, "
                               "Nphi=256, m=0.1, Kerr-Schild M=1, CFL 0.25, %s outer boundary" % (order, args.bc),
                   "grid_points": CFG["Nx"] * (CFG["Ny"] - 1) * (CFG["Nz"] - 1), "scheme": "rk4",
                   "operators": args.operators, "fields": 5, "state_finite": head["state_finite"],
                   "tuning": head["tuning"]},
        "roofline": head["roofline"],
        "kernel_timing": head["kernel_timing"],
        "fill_ms_per_step": head["fill_ms_per_step"],
        "outer_column_ms_per_step": head["outer_column_ms_per_step"],
        "host_transfer": head["host_transfer"],
    }
    other = "ref2" if args.operators == "fd4" else "fd4"
    if args.other_steps > 0:
        print("bench: the %s operator set beside it" % other, file=sys.stderr, flush=True)
        out["plain_" + other] = plain(other, args.other_steps, 5)
    if args.pirk2_steps > 0:
        print("bench: the reference's PIRK2 + 2nd-order step", file=sys.stderr, flush=True)
        out["plain_ref2_pirk2"] = plain("ref2", args.pirk2_steps, 5, kc=args.pirk2_kc, tj=8, scheme="pirk2")
    if args.refined_steps > 0:
        out["refined_pirk"] = [refined_workload(sch, args.refined_steps) for sch in ("pirk2", "pirk3", "rk3")]
    if args.ic > 0:
        print("bench: initial data, ncart %d" % args.ic, file=sys.stderr, flush=True)
        out["initial_data"] = initial_data(args.ic)
    # the batched and serving legs run the timed workload's operator set; a
    # grid run alone uses the launch shape its own autotuner picks (PLAIN_SHAPES),
    # batched members the chunk the batch autotuner picks, served ones the
    # chunk the workgroup-fill rule gives for the requests active in each call;
    # theta rows per workgroup are 8 throughout
    shape = (8, 8)
    if args.batched > 0:
        print("bench: batched vs plain, %d members" % args.batched, file=sys.stderr, flush=True)
        out["batched"] = batched_vs_plain(args.batched, args.batched_steps, shape, args.operators)
        print("bench: batched vs plain, %d small members" % args.batched, file=sys.stderr, flush=True)
        small = batched_vs_plain(args.batched, 50, shape, args.operators, specs=small_specs(args.batched),
                                 spectral=False)
        small.pop("spectral", None)
        small["workload"] = ("the population's smallest class alone: %d members of Nr=256, (Ntheta,Nphi)=(16,32), "
                             "m~U[0,0.5], M~U[0.5,2], 50 RK4 steps" % args.batched)
        out["batched_small_members"] = small
    if args.serve > 0:
        print("bench: serving loop, %d requests" % args.serve, file=sys.stderr, flush=True)
        out["serving"] = serving(args.serve, shape, args.operators, "fill")
    if args.serve_small > 0:
        print("bench: serving loop, %d small requests" % args.serve_small, file=sys.stderr, flush=True)
        out["serving_small_requests"] = serving(
            args.serve_small, shape, args.operators, "fill", specs=small_specs(args.serve_small, seed=20261205),
            workload="Poisson(50/s) requests of the population's smallest class (Nr=256, (Ntheta,Nphi)=(16,32), "
                     "m~U[0,0.5], M~U[0.5,2]), steps ~U{500..5000}, chunks of 50 RK4 steps, final fields + "
                     "l<=8, 512-k spectrum (split-BF16)")
    if not args.no_cpu:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.operators)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
