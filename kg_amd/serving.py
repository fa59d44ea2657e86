"""Serving loop: evolution requests admitted against an HBM budget and
advanced together by continuous batching.

The reference runs one evolution per (MPI) job (main/evolve.f90:22-29); here
many requests share one GPU.  A request is an initial-data/grid draw of the
uneven batched workload plus a step count.  ``ServingLoop.run``:

- moves requests that have arrived into a FIFO queue;
- admits from the head of the queue while the request's device footprint
  fits the remaining HBM budget (and an optional cap on active requests):
  its grid, state and spectral tables become resident;
- advances every active request by one chunk of RK4 steps with one batched
  launch per stage (``kg_batch_step``): the chunk is ``chunk`` steps, or fewer
  when an active request needs fewer to finish, so requests that arrive
  while others are on the GPU join at the next chunk boundary;
- finishes a request as soon as its steps are done: its final fields and its
  spherical-Bessel spectrum (l <= lmax, k_n = kmax n / nk) are read back and
  its memory is released for the next admission.

A request's result is the same as running it alone: batched stepping is
bit-identical to stepping each grid by itself, and the spectrum is computed
by the same kernels on the same state.

``ServingLoop`` drives an engine object (``GpuEngine`` here); the admission
and batching logic is independent of the GPU.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from .batch import MemberSpec, uneven_specs


@dataclass
class Request:
    rid: int
    spec: MemberSpec
    nsteps: int
    arrival: float  # seconds after the start of the stream


@dataclass
class Result:
    rid: int
    nsteps: int
    points: int
    arrival: float
    admitted: float
    done: float
    fields: object = None    # final padded state, SoA [part][field][k][j][i]
    spectrum: object = None  # complex [lm][n]
    error: str = ""

    @property
    def latency(self) -> float:
        """time-to-result: completion minus arrival (queueing included)"""
        return self.done - self.arrival


@dataclass
class ServeStats:
    wall: float = 0.0
    updates: int = 0
    chunks: int = 0
    max_bytes: int = 0
    max_active: int = 0
    joined_running: int = 0          # admissions while other requests were active
    batch_sizes: list = field(default_factory=list)

    @property
    def updates_per_s(self) -> float:
        return self.updates / self.wall if self.wall > 0 else 0.0


def poisson_requests(n: int, rate: float = 50.0, seed: int = 20261203, steps=(500, 5000), specs=None):
    """n requests of the uneven batched workload (or of the grids `specs`)
    arriving as a Poisson stream of `rate` per second, step counts uniform in
    [steps[0], steps[1]]."""
    rng = np.random.default_rng(seed + 1)
    specs = uneven_specs(n, seed) if specs is None else list(specs)[:n]
    gaps = rng.exponential(1.0 / rate, size=n)
    arrivals = np.cumsum(gaps) - gaps[0]
    nsteps = rng.integers(steps[0], steps[1] + 1, size=n)
    return [Request(i, specs[i], int(nsteps[i]), float(arrivals[i])) for i in range(n)]


def percentile(values, q):
    """q-th percentile with linear interpolation between order statistics."""
    v = np.sort(np.asarray(values, dtype=np.float64))
    if len(v) == 0:
        return float("nan")
    return float(np.percentile(v, q))


class ServingLoop:
    def __init__(self, engine, hbm_budget: int, chunk: int = 50, max_active: int = 1 << 30,
                 clock=time.perf_counter, sleep=time.sleep):
        if chunk < 1:
            raise ValueError("chunk must be >= 1")
        self.engine = engine
        self.budget = int(hbm_budget)
        self.chunk = int(chunk)
        self.max_active = int(max_active)
        self.clock = clock
        self.sleep = sleep

    def run(self, requests, keep_fields: bool = True):
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        queue = []
        active = []  # [request, member, bytes, remaining, admitted_at]
        results = {}
        used = 0
        stats = ServeStats()
        t0 = self.clock()
        pi = 0
        while pi < len(pending) or queue or active:
            now = self.clock() - t0
            while pi < len(pending) and pending[pi].arrival <= now:
                queue.append(pending[pi])
                pi += 1
            # admission, FIFO, against the HBM budget
            while queue and len(active) < self.max_active:
                req = queue[0]
                need = self.engine.bytes_for(req.spec)
                if need > self.budget:
                    queue.pop(0)
                    t = self.clock() - t0
                    results[req.rid] = Result(req.rid, req.nsteps, req.spec.points, req.arrival, t, t,
                                              error="needs %d bytes, budget %d" % (need, self.budget))
                    continue
                if used + need > self.budget:
                    break
                queue.pop(0)
                member = self.engine.admit(req)
                if active:
                    stats.joined_running += 1
                active.append([req, member, need, req.nsteps, self.clock() - t0])
                used += need
                stats.max_bytes = max(stats.max_bytes, used)
            stats.max_active = max(stats.max_active, len(active))
            if not active:
                if pi < len(pending):  # idle until the next arrival
                    self.sleep(max(0.0, pending[pi].arrival - (self.clock() - t0)))
                continue
            n = min(self.chunk, min(a[3] for a in active))
            self.engine.step([a[1] for a in active], n)
            stats.chunks += 1
            stats.batch_sizes.append(len(active))
            stats.updates += n * sum(a[0].spec.points for a in active)
            still = []
            for a in active:
                a[3] -= n
                if a[3] > 0:
                    still.append(a)
                    continue
                req, member = a[0], a[1]
                fields, spectrum = self.engine.finish(member, keep_fields)
                t = self.clock() - t0
                results[req.rid] = Result(req.rid, req.nsteps, req.spec.points, req.arrival, a[4], t,
                                          fields, spectrum)
                self.engine.release(member)
                used -= a[2]
            active = still
        stats.wall = self.clock() - t0
        return [results[r.rid] for r in requests], stats


class GpuEngine:
    """Requests as resident GPU grids (``LinesMethod``) with their spectral
    projector; ``step`` is one batched launch per RK4 stage."""

    def __init__(self, lmax: int = 8, nk: int = 512, kmax: float = 4.0, CFL: float = 0.25, tuning=None,
                 on_device: bool = True, operators: str = "ref2", batch_kc=None, fill_workgroups: int = 1024,
                 scheme: str = "rk4"):
        """tuning: (kc, tj) of a request run alone; batch_kc: the phi chunk of
        requests advanced together (defaults to tuning's), or "fill" to pick it
        per chunk from the active set (``autotune.fill_kc`` with
        ``fill_workgroups`` full-length chunks per stage)."""
        self.lmax, self.nk, self.kmax, self.CFL = lmax, nk, kmax, CFL
        self.tuning = tuning
        self.batch_kc = batch_kc
        self.fill_workgroups = int(fill_workgroups)
        self.scheme = scheme  # every request runs this integrator
        self.kc_chunks = {}  # phi chunk -> batched calls that used it
        self.on_device = on_device
        self.operators = operators

    def bytes_for(self, spec: MemberSpec) -> int:
        """Device footprint: four state buffers (RK4's u, next state / k1, k2, k3;
        the other schemes hold fewer, so this bounds them too),
        the metric tables and the spectral tables and scratch."""
        ni, nj, nk = spec.Nr + 5, spec.Ntheta + 1 + 5, spec.Nphi + 2
        ld = (ni + 15) // 16 * 16
        state = 4 * 5 * ld * nj * nk * 8
        tables = 10 * ld * nj * 8
        nlm = (self.lmax + 1) ** 2
        mc = 9 if self.lmax + 1 <= 9 else 17
        spectral = ((self.lmax + 1) * self.nk * spec.Nr + (spec.Ntheta + 1) * 2 * mc * spec.Nr
                    + 4 * nlm * spec.Nr + 2 * nlm * self.nk) * 8
        return int(1.02 * (state + tables + spectral)) + (4 << 20)

    def admit(self, req: Request):
        from .batch import make_member
        from .spectral import SpectralProjector

        lm = make_member(req.spec, CFL=self.CFL, on_device=self.on_device, operators=self.operators,
                         scheme=self.scheme)
        if self.tuning:
            lm.set_tuning(*self.tuning)
        sp = SpectralProjector(lm, lmax=self.lmax, nk=self.nk, basis="uniform", kmax=self.kmax)
        return (lm, sp)

    def step(self, members, nsteps: int):
        from .batch import batch_step

        if self.batch_kc:
            tj = self.tuning[1] if self.tuning else 8
            kc = self.batch_kc
            if kc == "fill":
                from .autotune import fill_kc

                kc = fill_kc([m[0] for m in members], tj=tj, min_workgroups=self.fill_workgroups)
            self.kc_chunks[kc] = self.kc_chunks.get(kc, 0) + 1
            for m in members:
                m[0].set_tuning(kc, tj)  # a no-op when unchanged
        batch_step([m[0] for m in members], nsteps)

    def finish(self, member, keep_fields: bool = True):
        lm, sp = member
        spectrum = sp.radial(0, "split_bf16")
        fields = lm.u_soa() if keep_fields else None
        return fields, spectrum

    def release(self, member):
        lm, sp = member
        sp.close()
        lm.close()

    def run_alone(self, req: Request, keep_fields: bool = True, tuning=None):
        """The same request through the plain path: its own grid, kg_step graph
        replay (at `tuning` = (kc, tj) when given), then the same read-back."""
        member = self.admit(req)
        try:
            if tuning:
                member[0].set_tuning(*tuning)
            member[0].step(req.nsteps)
            return self.finish(member, keep_fields)
        finally:
            self.release(member)
