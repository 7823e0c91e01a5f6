"""Serving loop: continuous batching of solve requests on one GPU.

Requests (a case, its initial state, an end time, an arrival time) queue up
as they arrive.  Between device rounds the loop
  * retires every case that reached its end time: the result is copied to
    the host, its latency (result on host - arrival) recorded, its slot and
    device memory released;
  * admits queued requests while a slot is free and the request's device
    memory (state + scratch, 46 FP64 planes of its padded grid) fits in the
    HBM budget -- the new cases join the running batch;
  * advances every resident case by one round of `round_steps` steps (one
    hipGraph replay).
Cases therefore enter and leave while others are running: their work
overlaps on the device in the same persistent launches.  A case's result
does not depend on what else is in the batch (tests/test_gpu_serve.py).
"""
from __future__ import annotations

import dataclasses
import heapq
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import cases
from .batch import Batch, Tuning
from .config import PREC_MIXED, SCHEME_MP5, CaseConfig

# device memory of a case in FP64 planes: 6 state + 2 x 6 stage planes + the 28
# face planes, which mixed precision keeps as 4 FP64 (energy) + 24 FP32 planes
PLANES_PER_CASE = {0: 6 + 12 + 28, 1: 6 + 12 + 4 + 12}


@dataclass(order=True)
class Request:
    """A solve request: case, initial state, end time, arrival time.  The
    initial state is a host array (U0) or, with U0 None, built on the device
    when the request is admitted (build(), e.g. cases.*_device)."""
    arrival: float
    rid: int = field(compare=False)
    cfg: CaseConfig = field(compare=False)
    U0: np.ndarray = field(compare=False, repr=False)
    t_end: float = field(compare=False)
    build: object = field(default=None, compare=False, repr=False)

    def state0(self):
        return self.U0 if self.U0 is not None else self.build()

    @property
    def nbytes(self) -> int:
        g = self.cfg.grid
        return PLANES_PER_CASE[self.cfg.precision] * g.nxg * g.nyg * 8


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    finished: float
    admitted_round: int
    finished_round: int
    state: np.ndarray
    t: float

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class Server:
    def __init__(self, scheme=SCHEME_MP5, precision: int = PREC_MIXED, max_slots: int = 64,
                 hbm_budget: float = 64e9, round_steps: int = 10, tuning: Tuning | None = None,
                 clock=time.perf_counter, sleep=time.sleep, batch=None, track_busy: bool = False):
        # `batch`: a Batch-like object to drive (the host logic's tests pass a stand-in)
        self.batch = batch if batch is not None else Batch(scheme, precision, max_slots)
        self.scheme, self.precision = scheme, precision
        self.max_slots = max_slots
        self.budget = float(hbm_budget)
        self.used = 0.0
        self.round_steps = round_steps
        self.clock = clock
        self.sleep = sleep
        self.peak_used = 0.0
        self.peak_resident = 0
        self.free = list(range(max_slots))
        self.slot_req: dict[int, tuple[Request, float, int]] = {}
        self.queue: list[Request] = []
        self.results: list[Result] = []
        self.rounds = 0
        # device-busy accounting: an event pair around every round's replay
        self.track_busy = track_busy
        self._events = []
        self.busy_s = None
        if tuning:
            self.batch.set_tuning(tuning)

    def submit(self, req: Request):
        if req.nbytes > self.budget:
            raise ValueError(f"request {req.rid} needs {req.nbytes / 1e9:.2f} GB > budget")
        heapq.heappush(self.queue, req)

    # one admission / retirement pass; returns True if the batch changed
    def _retire(self) -> bool:
        if not self.slot_req:
            return False
        t, _, done = self.batch.status()
        changed = False
        now = self.clock()
        for slot in [s for s in self.slot_req if done[s]]:
            req, admitted, round0 = self.slot_req.pop(slot)
            self.results.append(Result(req.rid, req.arrival, admitted, now, round0, self.rounds,
                                       self.batch.state(slot), float(t[slot])))
            self.batch.clear_case(slot)
            self.free.append(slot)
            self.used -= req.nbytes
            changed = True
        return changed

    def _admit(self, now: float) -> list[int]:
        new = []
        while self.queue and self.queue[0].arrival <= now and self.free:
            req = self.queue[0]
            if self.used + req.nbytes > self.budget:
                break  # first-come first-served: wait for memory
            heapq.heappop(self.queue)
            slot = self.free.pop(0)
            self.batch.set_case(slot, req.cfg, req.state0(), t0=0.0, t_end=req.t_end)
            self.used += req.nbytes
            self.slot_req[slot] = (req, now, self.rounds)
            new.append(slot)
        self.peak_used = max(self.peak_used, self.used)
        self.peak_resident = max(self.peak_resident, len(self.slot_req))
        return new

    def run(self, until_empty: bool = True, max_rounds: int = 10 ** 9, log=None):
        """Serve until every submitted request has finished.  With track_busy,
        busy_s is the device time of the rounds' replays (event pairs)."""
        t_start = self.clock()
        while self.rounds < max_rounds:
            now = self.clock()
            changed = self._retire()
            new = self._admit(now)
            if changed or new:
                self.batch.commit(graph_steps=self.round_steps if self.slot_req else 0)
                if new:
                    self.batch.prepare(new)
            if not self.slot_req:
                if not self.queue:
                    break
                # idle until the next arrival
                wait = self.queue[0].arrival - self.clock()
                if wait > 0:
                    self.sleep(min(wait, 0.05))
                continue
            if self.track_busy:
                import torch

                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                self.batch.step()
                e1.record()
                self._events.append((e0, e1))
            else:
                self.batch.step()
            self.rounds += 1
            if log and self.rounds % 200 == 0:
                log(f"  round {self.rounds}: resident {len(self.slot_req)}, queued {len(self.queue)}, "
                    f"done {len(self.results)}, HBM {self.used / 1e9:.2f} GB")
        elapsed = self.clock() - t_start
        if self.track_busy:
            import torch

            torch.cuda.synchronize()
            self.busy_s = sum(a.elapsed_time(b) for a, b in self._events) / 1e3
            self._events = []
        return elapsed

    def close(self):
        self.batch.close()


# ---------------------------------------------------------------------------
# The serving workload
# ---------------------------------------------------------------------------
T_END_DROPLET = 0.05
T_END_TRIPLE = 0.5


def make_request(rng, rid: int, arrival: float, scheme=SCHEME_MP5, precision=PREC_MIXED,
                 device_ic: bool = False):
    """Log-uniform NX in [128, 2048] (multiple of 32), NY = clamp(NX/2, 128, 1024);
    half shock-droplet cases (droplets ~ area, at least one), half triple points.
    device_ic: the request carries its case recipe and its initial state is
    built in device memory at admission instead of on the host now."""
    nx = int(round(math.exp(rng.uniform(math.log(128), math.log(2048))) / 32.0)) * 32
    nx = min(max(nx, 128), 2048)
    ny = min(max(int(round(nx / 2 / 32.0)) * 32, 128), 1024)
    if rng.uniform() < 0.5:
        from .config import Grid

        grid = Grid(nx, ny, 0.0, 4.0, 0.0, 2.0)
        n = max(1, int(round(40 * nx * ny / (4096 * 2048))))
        drops = cases.seeded_droplets(rng, n, 0.05, 0.15, 0.9, 3.9, 0.05, 1.95, 4.0 * grid.dx,
                                      n_cavity=n // 4)
        cfg = CaseConfig(grid, cases.WATER_ND, scheme, precision, name=f"droplets_{nx}x{ny}")
        t_end = T_END_DROPLET
        if device_ic:
            return Request(arrival, rid, cfg, None, t_end,
                           build=lambda: cases.shock_droplets_device(grid, cases.WATER_ND, 2.4, 0.8, drops))
        U0 = cases.shock_droplets(grid, cases.WATER_ND, 2.4, 0.8, drops)
    else:
        t_end = T_END_TRIPLE
        grid = cases.triple_point_grid(nx, ny)
        cfg = CaseConfig(grid, cases.TRIPLE_EOS, scheme, precision, name=f"triple_point_{nx}x{ny}")
        if device_ic:
            return Request(arrival, rid, cfg, None, t_end,
                           build=lambda: cases.triple_point_device(nx, ny, scheme=scheme)[1])
        U0 = cases.triple_point(nx, ny, scheme=scheme)[1]
    return Request(arrival, rid, cfg, U0, t_end)


def poisson_workload(n: int, rate: float, seed: int = 5, t0: float = 0.0, **kw):
    rng = np.random.default_rng(seed)
    t = t0
    out = []
    for k in range(n):
        t += rng.exponential(1.0 / rate)
        out.append(make_request(rng, k, t, **kw))
    return out


def with_arrivals(reqs, rate: float | None, base: float = 0.0, precision: int | None = None):
    """Copies of `reqs` whose arrivals are rescaled to a Poisson rate (`rate`
    requests/s; the workload was drawn at rate 1) or all at `base` (rate None)."""
    out = []
    for r in reqs:
        cfg = r.cfg if precision is None else dataclasses.replace(r.cfg, precision=precision)
        arrival = base if rate is None else base + r.arrival / rate
        out.append(dataclasses.replace(r, arrival=arrival, cfg=cfg))
    return out


def _stats(results, elapsed: float) -> dict:
    lat = np.array([r.latency for r in results])
    return {
        "completed": int(len(lat)),
        "elapsed_s": float(elapsed),
        "cases_per_hour": float(len(lat) / elapsed * 3600.0),
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "mean_latency_s": float(lat.mean()),
    }


def serve_plain(reqs, clock=time.perf_counter, keep=None, log=None):
    """The plain solver serving a queue: one request at a time in arrival order,
    each with its own Solver (set_state, prepare, the reference's time loop to
    t_end, result copied to the host).  `reqs` arrivals are offsets from the
    start.  `keep(rid, state)` receives each result."""
    import torch

    from .solver import Solver

    order = sorted(reqs)
    t0 = clock()
    results = []
    for r in order:
        now = clock() - t0
        if now < r.arrival:
            time.sleep(r.arrival - now)
        start = clock() - t0
        s = Solver(r.cfg)
        s.set_state(r.state0())
        s.prepare()
        s.advance_to(r.t_end, log=None)
        U = s.state()
        torch.cuda.synchronize()
        s.close()
        done = clock() - t0
        if keep:
            keep(r.rid, U)
        results.append(Result(r.rid, r.arrival, start, done, 0, 0, None, s.time))
        if log and len(results) % 25 == 0:
            log(f"plain: {len(results)} of {len(order)} served, {done:.1f} s")
    return results, _stats(results, clock() - t0)


def serve_batched(reqs, scheme=SCHEME_MP5, precision=PREC_MIXED, max_slots: int = 64,
                  hbm_budget: float = 64e9, round_steps: int = 10, tuning: Tuning | None = None,
                  log=None, track_busy: bool = False):
    """Run `reqs` (arrivals as offsets from the start) through a Server."""
    srv = Server(scheme=scheme, precision=precision, max_slots=max_slots, hbm_budget=hbm_budget,
                 round_steps=round_steps, tuning=tuning or Tuning(1, 3, 2, 0),
                 track_busy=track_busy)
    base = srv.clock()
    for r in reqs:
        srv.submit(dataclasses.replace(r, arrival=base + r.arrival))
    elapsed = srv.run(log=log)
    for r in srv.results:
        r.arrival -= base
        r.admitted -= base
        r.finished -= base
    st = _stats(srv.results, elapsed)
    st.update(rounds=srv.rounds, peak_resident=srv.peak_resident,
              peak_hbm_gb=round(srv.peak_used / 1e9, 3))
    if track_busy:
        st.update(device_busy_s=srv.busy_s, device_busy_frac=srv.busy_s / elapsed)
    results = srv.results
    srv.close()
    return results, st


def sweep(n: int = 200, loads=(0.3, 0.5, 0.7, 0.9, 1.2), seed: int = 5, log=None) -> dict:
    """Latency against offered load: the server's capacity from a saturated run
    (all requests queued at t=0), then Poisson arrivals at each load x capacity.
    Initial states are built on the device at admission."""
    reqs = poisson_workload(n, 1.0, seed, device_ic=True)
    warm = [r for r in reqs if r.cfg.grid.nx <= 192][:2]
    serve_batched(with_arrivals(warm, None))
    _, sat = serve_batched(with_arrivals(reqs, None))
    cap = sat["completed"] / sat["elapsed_s"]
    if log:
        log(f"capacity {cap:.3f} requests/s ({sat['cases_per_hour']:.0f} cases/hour)")
    rows = []
    for load in loads:
        rate = load * cap
        _, st = serve_batched(with_arrivals(reqs, rate))
        st.update(load=load, rate_per_s=rate)
        rows.append(st)
        if log:
            log(f"load {load:.2f}: p50 {st['p50_latency_s']:.3f} s  p99 {st['p99_latency_s']:.3f} s  "
                f"{st['cases_per_hour']:.0f} cases/hour")
    return {"requests": n, "seed": seed, "capacity_per_s": cap, "saturated": sat, "loads": rows}


def main(argv=None) -> int:
    """python -m wamp.serve [--requests N] [--loads 0.3,0.5,...]: the load sweep as JSON."""
    import argparse
    import json
    import sys

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--loads", type=str, default="0.3,0.5,0.7,0.9,1.2")
    ap.add_argument("--seed", type=int, default=5)
    args = ap.parse_args(argv)
    loads = [float(x) for x in args.loads.split(",") if x.strip()]
    out = sweep(args.requests, loads, args.seed, log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    import sys

    sys.exit(main())
