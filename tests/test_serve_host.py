"""Host logic of the serving loop, with a fake engine (CPU only).

Checks admission by footprint (backfill past a head request that does not fit
yet), that every request completes with its own step count, that cases join
and leave between rounds, and the latency bookkeeping."""
import numpy as np

from spade_amd.serve import Request, Server, poisson_requests, summarize
from spade_amd import CaseSpec


class FakeEngine:
    def __init__(self, blocks):
        self.free = blocks
        self.cases = {}
        self.next = 0
        self.log = []

    def fits(self, spec):
        return spec.blocks <= self.free

    def add(self, spec):
        assert spec.blocks <= self.free
        self.free -= spec.blocks
        cid = self.next
        self.next += 1
        self.cases[cid] = [spec, 0]
        self.log.append(("add", cid))
        return cid

    def step(self, k):
        assert self.cases
        for c in self.cases.values():
            c[1] += k
        self.log.append(("step", k, tuple(sorted(self.cases))))

    def download(self, cid):
        return np.full(3, float(self.cases[cid][1]))

    def diag(self, cid):
        n = self.cases[cid][1]
        return np.arange(n, dtype=float), np.zeros(n)

    def remove(self, cid):
        spec, _ = self.cases.pop(cid)
        self.free += spec.blocks
        self.log.append(("remove", cid))


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 0.001
        return self.t

    def sleep(self, dt):
        self.t += dt


def test_every_request_completes_with_its_steps():
    eng = FakeEngine(blocks=64)
    clk = FakeClock()
    reqs = [Request(i, CaseSpec(cells=(32, 32, 32), steps=5 + 3 * i), arrival=0.0) for i in range(6)]
    res = Server(eng, max_chunk=4, clock=clk, sleep=clk.sleep).run(reqs)
    assert sorted(r.request.rid for r in res) == list(range(6))
    for r in res:
        assert r.state[0] == r.request.spec.steps
        assert len(r.ke) == r.request.spec.steps
        assert r.latency > 0
    assert eng.free == 64 and not eng.cases


def test_backfill_admits_small_request_past_blocked_head():
    eng = FakeEngine(blocks=64)
    clk = FakeClock()
    big = CaseSpec(cells=(64, 64, 64), steps=8)       # 64 blocks: the whole pool
    small = CaseSpec(cells=(32, 32, 32), steps=8)     # 8 blocks
    reqs = [Request(0, small, 0.0), Request(1, big, 0.0), Request(2, small, 0.0)]
    Server(eng, max_chunk=8, clock=clk, sleep=clk.sleep).run(reqs)
    # case ids follow admission order: request 0 -> id 0, request 2 -> id 1, request 1 -> id 2
    first_step = [e for e in eng.log if e[0] == "step"][0]
    assert first_step[2] == (0, 1)                       # both small requests run first
    assert eng.log.index(("add", 2)) > eng.log.index(("remove", 0))   # the big one waits for the pool


def test_cases_join_running_batch():
    eng = FakeEngine(blocks=1000)
    clk = FakeClock()
    reqs = [Request(0, CaseSpec(cells=(32, 32, 32), steps=40), 0.0),
            Request(1, CaseSpec(cells=(32, 32, 32), steps=4), 0.010)]
    res = Server(eng, max_chunk=2, clock=clk, sleep=clk.sleep).run(reqs)
    steps = [e for e in eng.log if e[0] == "step"]
    assert any(len(s[2]) == 2 for s in steps)          # both resident in one round
    by_id = {r.request.rid: r for r in res}
    assert by_id[1].completed < by_id[0].completed       # the short late case leaves first


def test_poisson_arrivals_and_summary():
    reqs = poisson_requests(50, rate=100.0, make_spec=lambda rng: CaseSpec(steps=int(rng.integers(1, 5))), seed=3)
    t = [r.arrival for r in reqs]
    assert all(b >= a for a, b in zip(t, t[1:]))
    assert 0.2 < t[-1] < 1.0                              # 50 arrivals at 100/s
    eng = FakeEngine(blocks=10**6)
    clk = FakeClock()
    srv = Server(eng, clock=clk, sleep=clk.sleep)
    res = srv.run(reqs)
    s = summarize(res, srv.makespan, srv.cell_updates)
    assert s["requests"] == 50 and s["p99_latency_s"] >= s["p50_latency_s"] > 0


class DivergingEngine(FakeEngine):
    """A fake engine whose case `bad_cid` has a non-finite kinetic energy from step `at` on."""

    def __init__(self, blocks, bad_cid, at):
        super().__init__(blocks)
        self.bad_cid, self.at = bad_cid, at

    def status(self, cid):
        return self.at if cid == self.bad_cid and self.cases[cid][1] > self.at else -1


def test_diverged_case_retired_with_error_others_complete():
    eng = DivergingEngine(blocks=64, bad_cid=1, at=3)
    clk = FakeClock()
    reqs = [Request(i, CaseSpec(cells=(32, 32, 32), steps=20), arrival=0.0) for i in range(2)]
    reqs.append(Request(2, CaseSpec(cells=(32, 32, 32), steps=20), arrival=0.05))   # keeps rounds short
    res = Server(eng, max_chunk=2, clock=clk, sleep=clk.sleep).run(reqs)
    by = {r.request.rid: r for r in res}
    assert by[1].error is not None and "step 3" in by[1].error
    assert by[1].state[0] < 20                         # retired before its step count
    for rid in (0, 2):
        assert by[rid].error is None and by[rid].state[0] == 20
    assert eng.free == 64 and not eng.cases


def test_head_request_not_starved_by_backfill():
    """A steady stream of small requests keeps a pool of 64 blocks partly busy;
    a request for the whole pool at the head of the queue is admitted once it
    has waited max_head_wait: backfill stops and the freed slots wait for it."""
    eng = FakeEngine(blocks=64)
    clk = FakeClock()
    small = CaseSpec(cells=(32, 32, 32), steps=6)       # 8 blocks
    big = CaseSpec(cells=(64, 64, 64), steps=6)         # 64 blocks
    reqs = [Request(0, small, 0.0), Request(1, big, 0.0005)]
    reqs += [Request(2 + i, small, 0.001 + 0.002 * i) for i in range(200)]
    srv = Server(eng, max_chunk=2, clock=clk, sleep=clk.sleep, max_head_wait=0.05)
    res = srv.run(reqs)
    by = {r.request.rid: r for r in res}
    assert len(res) == 202
    later = [r for r in res if r.request.rid >= 2 and r.admitted > by[1].admitted]
    assert later, "the big request must be admitted while small ones still arrive"
    assert by[1].admitted - by[1].request.arrival < 0.05 + 0.05
