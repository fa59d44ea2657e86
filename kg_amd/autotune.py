"""Launch-shape autotuner for the stage kernel.

The stage kernel's workgroup covers 64 r-points x `tj` theta rows and marches
over `kc` phi planes.  The best shape depends on the grid (how many
workgroups fill the 256 CUs, how much halo each tile re-reads), so it is
measured: every candidate runs a few steps on the caller's own handle with
HIP events around each stage kernel, and the fastest is kept.  Results do
not depend on the shape (the kernel computes the same expressions in the
same order for any tiling), so tuning can run on live state: the state is
saved before and restored after.
"""
from __future__ import annotations

import time

import numpy as np

CANDIDATES = [(kc, tj) for tj in (4, 8) for kc in (4, 8, 16, 32, 64)]


def autotune(lm, candidates=CANDIDATES, steps=2, verbose=False):
    """Pick (kc, tj) for LinesMethod `lm`; returns (best, {shape: ms per stage kernel})."""
    if getattr(lm, "operators", "ref2") == "fd4":
        # the fd4 stages fix their theta rows per workgroup: only kc varies
        candidates = [(kc, 8) for kc in sorted({c[0] for c in candidates})]
    saved = lm.u_soa()
    t_saved = lm.t
    aux = None
    if lm.desc.scheme in (0, 1):  # PIRK: the stale stage arrays carry over
        aux = [lm.u_stage_soa(0)] + ([lm.u_stage_soa(1)] if lm.desc.scheme == 1 else [])
    timings = {}
    for kc, tj in candidates:
        lm.set_tuning(kc, tj)
        lm.step(1)  # warm-up (graph capture)
        lm.profile(True)
        lm.step(steps)
        ms, n = lm.profile_read(0)
        lm.profile(False)
        timings[(kc, tj)] = ms / max(n, 1)
        if verbose:
            print("autotune kc=%d tj=%d: %.4f ms/stage" % (kc, tj, timings[(kc, tj)]), flush=True)
    best = min(timings, key=timings.get)
    lm.set_tuning(*best)
    lm.set_u_soa(saved)
    lm.t = t_saved
    if aux is not None:
        for w, arr in enumerate(aux):
            lm.set_u_stage_soa(w, arr)
    return best, timings


def autotune_batch(members, candidates=(8, 16, 32, 64), steps=2, tj=8):
    """Pick one phi-chunk length for a batch (every member gets it) by timing
    `steps` batched steps per candidate after one warm-up step.  The members
    advance (1 + steps) steps per candidate: tune before the work whose state
    matters, or on copies.  Returns (kc, {kc: seconds per batched step})."""
    from .batch import batch_step

    timings = {}
    for kc in candidates:
        for m in members:
            m.set_tuning(kc, tj)
        batch_step(members, 1)
        members[0].sync()
        t0 = time.perf_counter()
        batch_step(members, steps)
        members[0].sync()
        timings[kc] = (time.perf_counter() - t0) / steps
    best = min(timings, key=timings.get)
    for m in members:
        m.set_tuning(best, tj)
    return best, timings



def stage_tile_planes(lm, tj=8):
    """Phi planes marched by all workgroups of one RK stage of LinesMethod
    `lm`, summed: the stages of a step followed by a ghost fill compute
    i in [1, ni-2], j in [g+1, g+Ny-1] and every phi plane, in tiles of 64
    r-points by `tj` theta rows, once per real system."""
    nsys = 2 if lm.is_complex else 1
    return (-(-(lm.ni - 2) // 64)) * (-(-(lm.Ny - 1) // tj)) * lm.Nz * nsys


def fill_kc(members, candidates=(8, 16, 32), tj=8, min_workgroups=1024):
    """The longest phi chunk that still leaves `min_workgroups` full-length
    chunks of work per stage over the batch (4 per CU), else the shortest
    candidate.  Longer chunks re-read fewer phi edges; shorter ones keep the
    CUs busy when few or small grids are active.  On one MI355X: 32 planes for
    one 2048 x 128 x 256 grid (the plain autotuner's pick), and for the
    served batch of 8 uneven-population requests (16 and 32 served them at
    the one-at-a-time rate, a threshold of 16 chunks per CU, which chose 8
    planes there, 4 % below it); the batch autotuner takes the longest chunk
    for the 256-member population and for 256 grids of 256 x 16 x 32."""
    work = sum(stage_tile_planes(m, tj) for m in members)
    cands = sorted(candidates)
    for kc in reversed(cands):
        if work >= kc * min_workgroups:
            return kc
    return cands[0]
