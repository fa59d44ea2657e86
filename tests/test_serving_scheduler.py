"""Admission and continuous batching of the serving loop, with a fake engine
and a simulated clock (CPU only; the GPU engine is tested in
test_gpu_serving.py)."""
import numpy as np
import pytest

from kg_amd.batch import MemberSpec
from kg_amd.serving import Request, ServingLoop, percentile, poisson_requests


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += dt


class FakeEngine:
    """Each step costs 1e-9 s per grid-point update; records every call."""

    def __init__(self, clock, footprint=lambda spec: spec.points * 100):
        self.clock = clock
        self.footprint = footprint
        self.steps_done = {}
        self.log = []
        self.live = set()

    def bytes_for(self, spec):
        return self.footprint(spec)

    def admit(self, req):
        self.live.add(req.rid)
        self.steps_done[req.rid] = 0
        self.log.append(("admit", req.rid, self.clock()))
        return req

    def step(self, members, n):
        assert all(m.rid in self.live for m in members)
        for m in members:
            self.steps_done[m.rid] += n
        self.clock.t += 1e-9 * n * sum(m.spec.points for m in members)
        self.log.append(("step", tuple(m.rid for m in members), n))

    def finish(self, member, keep_fields):
        return ("fields", member.rid), ("spectrum", member.rid)

    def release(self, member):
        self.live.remove(member.rid)


def _spec(nr, nt=16, npx=32):
    return MemberSpec(nr, nt, npx, 0.1, 1.0, 1.8, 1.8 + nr * 0.2, 5.0, 1.5, {}, 0)


def test_each_request_runs_exactly_its_steps_and_budget_holds():
    clock = FakeClock()
    eng = FakeEngine(clock)
    reqs = [Request(i, _spec(100 + 37 * i), 10 + 7 * i, 0.001 * i) for i in range(9)]
    budget = 3 * max(eng.bytes_for(r.spec) for r in reqs)
    loop = ServingLoop(eng, budget, chunk=8, clock=clock, sleep=clock.sleep)
    results, stats = loop.run(reqs)
    assert [r.rid for r in results] == list(range(9))
    for r, q in zip(results, reqs):
        assert eng.steps_done[q.rid] == q.nsteps
        assert r.fields == ("fields", q.rid) and not r.error
        assert r.arrival <= r.admitted <= r.done
    assert stats.max_bytes <= budget
    assert stats.max_active >= 2  # several requests shared the GPU
    assert stats.updates == sum(q.nsteps * q.spec.points for q in reqs)
    assert not eng.live  # everything released


def test_continuous_batching_admits_while_others_run():
    clock = FakeClock()
    eng = FakeEngine(clock)
    # a long request and a stream of short ones that arrive while it runs
    reqs = [Request(0, _spec(2000), 400, 0.0)] + [Request(i, _spec(300), 20, 0.0005 * i) for i in range(1, 6)]
    loop = ServingLoop(eng, hbm_budget=10 ** 12, chunk=50, clock=clock, sleep=clock.sleep)
    results, stats = loop.run(reqs)
    assert stats.joined_running >= 5
    long_done = results[0].done
    # the short requests finish long before the long one
    assert all(r.done < long_done for r in results[1:])
    # chunks shrink to what a request still needs: no request is over-stepped
    assert all(eng.steps_done[q.rid] == q.nsteps for q in reqs)


def test_fifo_admission_and_queueing_under_a_tight_budget():
    clock = FakeClock()
    eng = FakeEngine(clock, footprint=lambda spec: 1000)
    reqs = [Request(i, _spec(200), 30, 0.0) for i in range(5)]
    loop = ServingLoop(eng, hbm_budget=2000, chunk=10, clock=clock, sleep=clock.sleep)
    results, stats = loop.run(reqs)
    admits = [e[1] for e in eng.log if e[0] == "admit"]
    assert admits == [0, 1, 2, 3, 4]
    assert stats.max_active == 2
    assert results[2].admitted >= results[0].done  # waited for memory
    assert results[4].latency > results[0].latency


def test_oversized_request_is_rejected_not_deadlocked():
    clock = FakeClock()
    eng = FakeEngine(clock)
    reqs = [Request(0, _spec(5000), 5, 0.0), Request(1, _spec(100), 5, 0.0)]
    loop = ServingLoop(eng, hbm_budget=eng.bytes_for(_spec(1000)), chunk=5, clock=clock, sleep=clock.sleep)
    results, _ = loop.run(reqs)
    assert results[0].error and results[0].fields is None
    assert not results[1].error


def test_idle_gpu_waits_for_the_next_arrival():
    clock = FakeClock()
    eng = FakeEngine(clock)
    reqs = [Request(0, _spec(100), 5, 0.0), Request(1, _spec(100), 5, 3.0)]
    loop = ServingLoop(eng, hbm_budget=10 ** 12, chunk=5, clock=clock, sleep=clock.sleep)
    results, stats = loop.run(reqs)
    assert results[1].admitted >= 3.0
    assert stats.wall >= 3.0


def test_poisson_stream_and_percentiles():
    reqs = poisson_requests(200, rate=50.0, seed=5)
    gaps = np.diff([r.arrival for r in reqs])
    assert reqs[0].arrival == 0.0 and np.all(gaps >= 0)
    assert 0.014 < gaps.mean() < 0.027  # mean 1/50 s
    assert all(500 <= r.nsteps <= 5000 for r in reqs)
    assert poisson_requests(10, seed=5)[3].nsteps == poisson_requests(10, seed=5)[3].nsteps
    assert percentile([1, 2, 3, 4], 50) == pytest.approx(2.5)
    assert percentile(list(range(101)), 99) == pytest.approx(99.0)


def test_fill_rule_phi_chunks():
    """The workgroup-fill rule: the longest phi chunk leaving 1024 full-length
    chunks per stage (4 per CU): one 2048x128x256 grid takes 32 (the plain
    autotuner's pick), 256 grids of 256x16x32 (33 phi planes each) the
    longest candidate, and its fallback."""
    from types import SimpleNamespace

    from kg_amd.autotune import fill_kc, stage_tile_planes

    def grid(Nr, Nt, Np, cx=False):
        return SimpleNamespace(ni=Nr + 5, Ny=Nt + 1, Nz=Np + 1, is_complex=cx)

    big = grid(2048, 128, 256)
    assert stage_tile_planes(big) == 33 * 16 * 257
    assert stage_tile_planes(grid(2048, 128, 256, True)) == 2 * 33 * 16 * 257
    assert fill_kc([big]) == 32
    assert fill_kc([grid(256, 16, 32)] * 256) == 32
    assert fill_kc([grid(256, 16, 32)] * 60) == 16  # 19,800 tile planes
    assert fill_kc([grid(256, 16, 32)]) == 8  # nothing fills: the shortest
    assert fill_kc([big] * 4, candidates=(8, 16, 32, 64)) == 64
    assert fill_kc([grid(24, 8, 16)], min_workgroups=1) == 16  # 17 planes: a 32 chunk would leave < 1


def test_poisson_requests_take_a_grid_list():
    """A stream of given grids (the small-request serving workload): the
    requests carry those grids in order, with the stream's arrivals and steps."""
    specs = [MemberSpec(256, 16, 32, 0.1 * q, 1.0, 1.8, 60.0, 15.0, 1.5, {(0, 0): 1.0}, q) for q in range(5)]
    reqs = poisson_requests(5, specs=specs, seed=9)
    assert [r.spec for r in reqs] == specs
    base = poisson_requests(5, seed=9)
    assert [r.arrival for r in reqs] == [r.arrival for r in base]
    assert [r.nsteps for r in reqs] == [r.nsteps for r in base]
