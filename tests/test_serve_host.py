"""Host logic of the continuous-batching server (wamp/serve.py) on the CPU: a
stand-in batch with the Batch interface finishes each case after a fixed
number of steps, so admission by HBM budget, first-come first-served order,
slot reuse, retirement, idling and the latency bookkeeping are exercised
without a GPU."""
import dataclasses

import numpy as np
import pytest

from wamp import serve
from wamp.config import PREC_MIXED, SCHEME_MP5, CaseConfig, Grid
from wamp.cases import WATER_ND


class FakeBatch:
    """Each case finishes `need[rid]` steps after its prepare; records calls."""

    def __init__(self, max_slots, need):
        self.max_slots = max_slots
        self.need = need
        self.slot_rid = {}
        self.left = {}
        self.t = np.zeros(max_slots)
        self.commits = 0
        self.graph_steps = 0
        self.log = []

    def set_tuning(self, tuning):
        pass

    def set_case(self, slot, cfg, U0, t0=0.0, t_end=0.0):
        assert slot not in self.slot_rid, "slot reused before it was cleared"
        self.slot_rid[slot] = int(U0[0, 0, 0])
        self.t[slot] = t0
        self.tend = getattr(self, "tend", {})
        self.tend[slot] = t_end

    def clear_case(self, slot):
        del self.slot_rid[slot]
        self.left.pop(slot, None)

    def commit(self, graph_steps=0):
        self.commits += 1
        self.graph_steps = graph_steps

    def prepare(self, slots=None):
        for s in slots:
            self.left[s] = self.need[self.slot_rid[s]]

    def step(self):
        n = self.graph_steps or 1
        for s in list(self.left):
            self.left[s] = max(0, self.left[s] - n)
            if self.left[s] == 0:
                self.t[s] = self.tend[s]
        self.log.append(sorted(self.slot_rid.values()))
        return n

    def status(self):
        done = np.zeros(self.max_slots, dtype=bool)
        for s in self.slot_rid:
            done[s] = self.left.get(s, 1) == 0
        return self.t.copy(), np.zeros(self.max_slots), done

    def state(self, slot):
        return np.full((6, 4, 4), float(self.slot_rid[slot]))

    def close(self):
        pass


class RoundClock:
    def __init__(self):
        self.srv = None
        self.idle = 0.0

    def __call__(self):
        return float(self.srv.rounds) + self.idle

    def sleep(self, d):
        self.idle += d


def _req(rid, arrival, nx, ny, t_end=0.1):
    cfg = CaseConfig(Grid(nx, ny, 0.0, 1.0, 0.0, 1.0), WATER_ND, SCHEME_MP5, PREC_MIXED)
    U0 = np.full((6, 4, 4), float(rid))  # the stand-in reads the request id from the state
    return serve.Request(arrival, rid, cfg, U0, t_end)


def _run(reqs, need, max_slots, budget, round_steps=2):
    fb = FakeBatch(max_slots, need)
    clock = RoundClock()
    srv = serve.Server(SCHEME_MP5, PREC_MIXED, max_slots=max_slots, hbm_budget=budget,
                       round_steps=round_steps, clock=clock, sleep=clock.sleep, batch=fb)
    clock.srv = srv
    for r in reqs:
        srv.submit(r)
    srv.run()
    return srv, fb


def test_request_bytes_follow_the_padded_grid():
    """State, two stage buffers and the face planes: 28 FP64 planes, or in mixed
    precision 4 FP64 energy planes + 24 FP32 planes."""
    import dataclasses

    r = _req(0, 0.0, 100, 60)
    g = r.cfg.grid
    planes = {0: 6 + 12 + 28, 1: 6 + 12 + 4 + 24 / 2}
    for prec, n in planes.items():
        rp = dataclasses.replace(r, cfg=dataclasses.replace(r.cfg, precision=prec))
        assert rp.nbytes == n * g.nxg * g.nyg * 8


def test_all_requests_finish_with_their_own_state_and_end_time():
    reqs = [_req(k, float(k), 32, 32) for k in range(6)]
    need = {k: 3 + 2 * k for k in range(6)}
    srv, fb = _run(reqs, need, max_slots=3, budget=1e12)
    assert sorted(r.rid for r in srv.results) == list(range(6))
    for r in srv.results:
        assert np.all(r.state == r.rid)  # the retired state is the request's own
        assert r.t == pytest.approx(0.1)
        assert r.finished >= r.admitted >= r.arrival
    assert srv.peak_resident <= 3
    assert not fb.slot_rid  # every slot was released


def test_admission_respects_the_budget_and_is_first_come_first_served():
    big, small = _req(0, 0.0, 256, 256), _req(1, 0.0, 32, 32)
    later_small = _req(2, 1.0, 32, 32)
    budget = big.nbytes + small.nbytes // 2  # the big one and nothing else fits
    srv, fb = _run([big, small, later_small], {0: 10, 1: 2, 2: 2}, max_slots=4, budget=budget)
    assert srv.peak_used <= budget
    by = {r.rid: r for r in srv.results}
    # request 1 may not overtake the waiting head of the queue, and waits for memory
    assert by[1].admitted_round >= by[0].finished_round
    assert by[2].admitted_round >= by[1].admitted_round
    for resident in fb.log:
        assert not (0 in resident and 1 in resident)


def test_request_over_budget_is_refused():
    r = _req(0, 0.0, 256, 256)
    fb = FakeBatch(2, {0: 1})
    srv = serve.Server(SCHEME_MP5, PREC_MIXED, max_slots=2, hbm_budget=r.nbytes - 1, batch=fb)
    with pytest.raises(ValueError):
        srv.submit(r)


def test_idle_until_the_next_arrival():
    reqs = [_req(0, 0.0, 32, 32), _req(1, 500.0, 32, 32)]
    srv, fb = _run(reqs, {0: 2, 1: 2}, max_slots=2, budget=1e12)
    by = {r.rid: r for r in srv.results}
    assert by[1].admitted >= 500.0
    assert by[0].finished < 500.0
    assert srv.rounds < 10  # no rounds were spent while the batch was empty


def test_slots_are_reused_and_the_batch_recommitted_on_change():
    reqs = [_req(k, 0.0, 32, 32) for k in range(5)]
    srv, fb = _run(reqs, {k: 2 for k in range(5)}, max_slots=2, budget=1e12)
    assert len(srv.results) == 5
    assert srv.peak_resident == 2
    assert fb.commits >= 3  # admissions and retirements change the resident set


def test_workload_generator_is_seeded_and_in_range():
    a = serve.poisson_workload(12, 2.0, seed=9)
    b = serve.poisson_workload(12, 2.0, seed=9)
    assert [r.cfg.grid.nx for r in a] == [r.cfg.grid.nx for r in b]
    assert all(128 <= r.cfg.grid.nx <= 2048 and 128 <= r.cfg.grid.ny <= 1024 for r in a)
    arr = [r.arrival for r in a]
    assert arr == sorted(arr) and arr[0] > 0
    c = serve.with_arrivals(a, 4.0)
    assert [r.arrival for r in c] == pytest.approx([x / 4.0 for x in arr])
    d = serve.with_arrivals(a, None, precision=0)
    assert all(r.arrival == 0.0 and r.cfg.precision == 0 for r in d)
