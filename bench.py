#!/usr/bin/env python3
"""Throughput of the FP64 SSP-RK3 step on one MI355X: Gcell-updates/s.

Workload: the Mach-10 water-cylinder case, 4096x2048 cells on [0,4]x[0,2]
(water cylinder r=0.25 at (1,1), gamma=4.4, p_inf=6000 in water, air
gamma=1.4; Mach 10 shock at x=0.6), the M10 scheme (MUSCL + THINC, Wong face
and flux limiters), CFL 0.4, FP64.  Synthetic initial state built on the
host, resident in HBM before the timed region.

One step = one full SSP-RK3 step (3 stages of fluxes, 3-stage Wong limiter,
RK combine, repair, alpha clamp) plus the step's dt reduction.
value = NX*NY*steps / t / 1e9.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Also reported, for the dominant kernel (the face-flux kernel), its average
launch duration from hipEvents recorded around it inside the timed region,
and as cpu_baseline the CPU oracle on the host cores over a bounded sample;
cpu_device is the product's own CPU device (--cpu) on the same case.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "wave-multiphase_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 vector
VALU_SIMDS = 1024          # 256 CUs x 4 SIMDs
CLOCK_HZ = 2.4e9
ALG_BYTES_PER_CELL_UPDATE = 384.0  # FP64 state: 96 + 144 + 144 B over the 3 stages


def cpu_threads() -> int:
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def cpu_baseline(cfg, U0, steps: int):
    """CPU oracle (test infrastructure) on the same workload: `steps` steps, all threads."""
    from oracle.oracle import Oracle

    th = cpu_threads()
    orc = Oracle(cfg, threads=th)
    orc.set_state(U0)
    orc.prepare()
    dt = orc.get_dt()
    t0 = time.perf_counter()
    for _ in range(steps):
        orc.step(dt)
        dt = orc.get_dt()
    el = time.perf_counter() - t0
    orc.close()
    g = cfg.grid
    return {
        "value": g.nx * g.ny * steps / el / 1e9,
        "unit": "Gcell-updates/s",
        "cores": th,
        "kind": "port",
        "sample": f"{steps} SSP-RK3 step(s) of the same {g.nx}x{g.ny} case, C restatement "
                  f"of the reference kernels (OpenMP, {th} threads), {el:.1f} s",
    }


def cpu_device_leg(cfg, U0, steps: int):
    """The product's CPU device (--cpu: the FP64 step's own code on the host) on the
    same workload: `steps` steps, the threads the process may use."""
    from wamp.cpu import CpuSolver

    th = cpu_threads()
    s = CpuSolver(cfg, threads=th)
    s.set_state(U0)
    s.prepare()
    dt = s.get_dt()
    t0 = time.perf_counter()
    for _ in range(steps):
        dt = s.step_dt(dt)
    el = time.perf_counter() - t0
    s.close()
    g = cfg.grid
    return {
        "value": g.nx * g.ny * steps / el / 1e9,
        "unit": "Gcell-updates/s",
        "threads": th,
        "sample": f"{steps} SSP-RK3 step(s) of the same {g.nx}x{g.ny} case on the CPU device "
                  f"(wamp_cpu_step, {th} threads), {el:.1f} s",
    }


def mixed_leg(cfg, U0, warmup: int, steps: int):
    """Same case and step count in the mixed precision mode (FP32 face arithmetic)."""
    import dataclasses
    import ctypes as C

    import torch

    from wamp import _lib
    from wamp.config import PREC_MIXED
    from wamp.solver import Solver

    g = cfg.grid
    s = Solver(dataclasses.replace(cfg, precision=PREC_MIXED), device="cuda:0")
    s.set_state(U0)
    s.prepare()
    s.compute_dt()
    for _ in range(warmup):
        s.step_device()
    _lib.check(s.L.wamp_timing_enable(s.ctx, steps), "wamp_timing_enable")
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        s.step_device()
    e1.record()
    torch.cuda.synchronize()
    t_ms = e0.elapsed_time(e1)
    ms = (C.c_double * 2)()
    nl = (C.c_int64 * 2)()
    _lib.check(s.L.wamp_timing_read(s.ctx, ms, nl), "wamp_timing_read")
    s.close()
    return {
        "value": round(g.nx * g.ny * steps / (t_ms / 1e3) / 1e9, 6),
        "unit": "Gcell-updates/s",
        "ms_per_step": round(t_ms / steps, 4),
        "precision": "fp32 faces (energy flux fp64) + fp64 pressure/limiter/RK",
        "kernels_ms": {"k_faces": round(ms[0] / max(nl[0], 1), 4),
                       "k_limit_update": round(ms[1] / max(nl[1], 1), 4)},
    }


def threed_leg(n: int, warmup: int, steps: int):
    """The 3D extension on a shock-sphere case (2n x n x n cells, M10 scheme),
    FP64 and FP32-mixed, the mixed run's error against the FP64 run (same dt
    sequence) after the timed steps."""
    import dataclasses

    import torch

    from wamp import threed

    cfg, U0 = threed.shock_sphere(n)
    g = cfg.grid
    cells = g.nx * g.ny * g.nz
    res, states, dts = {}, {}, None
    for prec in (0, 1):
        s = threed.Solver3(dataclasses.replace(cfg, precision=prec), device="cuda:0")
        s.set_state(U0)
        s.prepare()
        if prec == 0:
            s.compute_dt()
            dts = []
        for k in range(warmup + steps):
            if k == warmup:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                torch.cuda.synchronize()
                e0.record()
            if prec == 0:
                dts.append(s.dt_buf[0].clone())
                s.step_device()
            else:
                s.dt_buf[0].copy_(dts[k])
                s.step_with_device_dt()
        e1.record()
        torch.cuda.synchronize()
        t_ms = e0.elapsed_time(e1)
        res[prec] = {"value": round(cells * steps / (t_ms / 1e3) / 1e9, 6),
                     "ms_per_step": round(t_ms / steps, 4),
                     "scratch_gb": round(s.scratch_bytes() / 1e9, 3),
                     "state_finite": bool(torch.isfinite(s.U).all().item())}
        states[prec] = s.state()
        s.close()
    err = threed.compare3(cfg, states[1], states[0])
    return dict(res[0], workload=cfg.name, scheme="m10", unit="Gcell-updates/s", steps=steps,
                mixed=dict(res[1], vs_fp64_after_steps={k: float(v) for k, v in err.items()
                                                        if k != "finite"}))


def threed_batched_leg(n_cases: int, steps: int, seed: int = 6):
    """Uneven 3D shock-sphere cases (2n x n x n, n from {8, 12, 16, 24, 32},
    Mach uniform in [1.5, 4], M10): Solver3 one case after another in FP64
    against one 3D batch (each kernel of a stage one launch over every case,
    a 10-step graph) in FP64 and in mixed precision.  Steps only are timed,
    setup apart; the FP64 batch must equal the plain runs bit for bit."""
    import dataclasses

    import numpy as np
    import torch

    from wamp import threed
    from wamp.config import SCHEME_M10

    rng = np.random.default_rng(seed)
    work = [threed.shock_sphere(int(rng.choice([8, 12, 16, 24, 32])), SCHEME_M10,
                                float(rng.uniform(1.5, 4.0))) for _ in range(n_cases)]
    cells = sum(c.grid.nx * c.grid.ny * c.grid.nz for c, _ in work)
    plain, t_plain = [], 0.0
    for cfg, U0 in work:
        s = threed.Solver3(cfg)
        s.set_state(U0)
        s.prepare()
        s.compute_dt()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            s.step_device()
        torch.cuda.synchronize()
        t_plain += time.perf_counter() - t0
        plain.append(s.state())
        s.close()
    res, states = {}, {}
    for prec in (0, 1):
        b = threed.Batch3(SCHEME_M10, prec, max_cases=n_cases)
        for k, (cfg, U0) in enumerate(work):
            b.set_case(k, dataclasses.replace(cfg, precision=prec), U0)
        b.commit(graph_steps=10 if steps % 10 == 0 else 1)
        b.prepare()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        done = 0
        while done < steps:
            done += b.step()
        torch.cuda.synchronize()
        res[prec] = time.perf_counter() - t0
        states[prec] = [b.state(k) for k in range(n_cases)]
        b.close()
    same = all(np.array_equal(a, p_) for a, p_ in zip(states[0], plain))
    err = {}
    for (cfg, _), a, p_ in zip(work, states[1], states[0]):
        for key, v in threed.compare3(cfg, a, p_).items():
            if key != "finite":
                err[key] = max(err.get(key, 0.0), float(v))
    cph = lambda t: round(n_cases / t * 3600.0, 1)
    return {"workload": f"{n_cases} shock-sphere cases, 2n x n x n with n in {{8, 12, 16, 24, 32}}, "
                        f"Mach U[1.5, 4], M10, seed {seed}",
            "cells": cells, "steps_per_case": steps, "timing": "steps only",
            "plain_fp64_cases_per_hour": cph(t_plain),
            "batched_fp64_cases_per_hour": cph(res[0]),
            "batched_mixed_cases_per_hour": cph(res[1]),
            "batching_speedup_fp64": round(t_plain / res[0], 3),
            "speedup": round(t_plain / res[1], 3),
            "batched_fp64_gcell_updates_per_s": round(cells * steps / res[0] / 1e9, 4),
            "batched_mixed_gcell_updates_per_s": round(cells * steps / res[1] / 1e9, 4),
            "plain_fp64_gcell_updates_per_s": round(cells * steps / t_plain / 1e9, 4),
            "batched_fp64_equals_plain_bitwise": same,
            "mixed_vs_fp64_max_rel_l1": err}


def _timed_batch(work, precision: int, steps: int):
    """One batch of `work` in `precision`.  Setup (case table, initial
    sequence, autotuning, graph capture and one warm-up replay from the
    initial states) and stepping are timed separately; returns (setup s,
    stepping s, tuning, graph steps, the batch with its final states)."""
    import dataclasses

    import torch

    from wamp.batch import Batch
    from wamp.config import SCHEME_MP5

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b = Batch(SCHEME_MP5, precision, max_cases=len(work))
    for k, (cfg, U0) in enumerate(work):
        b.set_case(k, dataclasses.replace(cfg, precision=precision), U0)
    b.commit()
    b.prepare()
    tuning = b.autotune(steps=2)
    gsteps = 10 if steps % 10 == 0 else 1
    b.capture(gsteps)
    snap = {k: v.clone() for k, v in b.states.items()}
    b.step()  # first replay (warm-up), then back to the initial states
    for k, v in snap.items():
        b.states[k].copy_(v)
    b.prepare()
    del snap
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    done = 0
    while done < steps:
        done += b.step()
    torch.cuda.synchronize()
    return t1 - t0, time.perf_counter() - t1, tuning, gsteps, b


def batched_leg(n_cases: int, steps: int, seed: int = 4):
    """64 seeded uneven shock-droplet cases, `steps` steps each: the plain FP64
    solver one case at a time versus one batch in mixed precision (autotuned,
    step graph-captured), and, to separate batching from precision, the same
    batch in FP64.  Every mode is timed over the same boundary: the steps
    alone (cases/hour), with the setup (plain: context, upload, initial
    sequence, first dt; batched: case table, initial sequence, autotuning,
    capture) timed separately and also reported included."""
    import numpy as np
    import torch

    from wamp import accuracy, cases
    from wamp.config import PREC_FP64, PREC_MIXED, SCHEME_MP5
    from wamp.solver import Solver

    rng = np.random.default_rng(seed)
    work = [cases.batch_case(rng, scheme=SCHEME_MP5) for _ in range(n_cases)]
    cells = sum(c.grid.nx * c.grid.ny for c, _ in work)
    # plain FP64, one case after another (each case's state resident on the GPU)
    plain_states = []
    t_setup = t_steps = 0.0
    for cfg, U0 in work:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s = Solver(cfg)
        s.set_state(U0)
        s.prepare()
        s.compute_dt()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            s.step_device()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        t_setup += t1 - t0
        t_steps += t2 - t1
        if len(plain_states) < 4:
            plain_states.append(s.state())
        s.close()
    setup64, t_b64, _, _, b = _timed_batch(work, PREC_FP64, steps)
    same64 = all(np.array_equal(b.state(k)[:, cfg.grid.interior()[0], cfg.grid.interior()[1]],
                                plain_states[k][:, cfg.grid.interior()[0], cfg.grid.interior()[1]])
                 for k, (cfg, _) in enumerate(work[:len(plain_states)]))
    b.close()
    setup_mx, t_batch, tuning, gsteps, b = _timed_batch(work, PREC_MIXED, steps)
    spot = []
    for k in range(min(4, n_cases)):
        cfg = work[k][0]
        spot.append(accuracy.compare(cfg, b.state(k), plain_states[k]))
    b.close()
    cph = lambda t: round(n_cases / t * 3600.0, 2)
    return {
        "workload": f"{n_cases} shock-droplet cases, grids from "
                    f"{sorted(set((c.grid.nx, c.grid.ny) for c, _ in work))}, seed {seed}",
        "cells": cells,
        "steps_per_case": steps,
        "timing": "steps only (setup timed separately, and included in *_incl_setup)",
        "plain_fp64_cases_per_hour": cph(t_steps),
        "batched_mixed_cases_per_hour": cph(t_batch),
        "speedup": round(t_steps / t_batch, 3),
        "batched_fp64_cases_per_hour": cph(t_b64),
        "batching_speedup_fp64": round(t_steps / t_b64, 3),
        "plain_fp64_gcell_updates_per_s": round(cells * steps / t_steps / 1e9, 4),
        "batched_mixed_gcell_updates_per_s": round(cells * steps / t_batch / 1e9, 4),
        "batched_fp64_gcell_updates_per_s": round(cells * steps / t_b64 / 1e9, 4),
        "setup_s": {"plain_fp64": round(t_setup, 3), "batched_fp64": round(setup64, 3),
                    "batched_mixed": round(setup_mx, 3)},
        "cases_per_hour_incl_setup": {"plain_fp64": cph(t_steps + t_setup),
                                      "batched_fp64": cph(t_b64 + setup64),
                                      "batched_mixed": cph(t_batch + setup_mx)},
        "speedup_incl_setup": round((t_steps + t_setup) / (t_batch + setup_mx), 3),
        "batched_fp64_equals_plain_bitwise": bool(same64),
        "tuning": list(tuning.as_tuple()),
        "graph_steps": gsteps,
        "spot_check_vs_fp64": {
            k: float(max(e[k] for e in spot)) for k in ("rho", "p", "alpha1", "centroid_shift_cells")
        },
    }


def serving_leg(n: int, seed: int = 5, load: float = 0.7):
    """Continuous batching of `n` requests (log-uniform grids 128..2048 wide; half
    shock-droplet cases to t=0.05, half triple points to t=0.5).

    plain      the plain FP64 solver serving the queue one request at a time
    saturated  the batched mixed server with every request queued at t=0
               (its cases/hour is the server's capacity)
    poisson    the batched mixed server under Poisson arrivals at `load` x capacity
    Results of both server runs are checked bitwise against each other (same
    cases, different mixes) and against the plain FP64 results (relative L1)."""
    import numpy as np

    from wamp import accuracy, serve
    from wamp.config import PREC_FP64

    reqs = serve.poisson_workload(n, 1.0, seed, device_ic=True)  # states built in HBM at admission
    warm = [r for r in reqs if r.cfg.grid.nx <= 192][:2]
    serve.serve_plain(serve.with_arrivals(warm, None, precision=PREC_FP64))
    serve.serve_batched(serve.with_arrivals(warm, None))
    plain_states = {}
    _, plain = serve.serve_plain(serve.with_arrivals(reqs, None, precision=PREC_FP64),
                                 keep=lambda rid, U: plain_states.__setitem__(rid, U), log=_progress)
    _progress(f"serving: plain fp64 {n} requests {plain['elapsed_s']:.1f} s")
    res_sat, sat = serve.serve_batched(serve.with_arrivals(reqs, None), log=_progress,
                                       track_busy=True)
    _progress(f"serving: batched saturated {sat['elapsed_s']:.1f} s")
    rate = load * sat["completed"] / sat["elapsed_s"]
    res_poi, poi = serve.serve_batched(serve.with_arrivals(reqs, rate), log=_progress)
    _progress(f"serving: batched poisson at {rate:.2f}/s {poi['elapsed_s']:.1f} s")
    by_rid = {r.rid: r for r in reqs}
    poi_states = {r.rid: r.state for r in res_poi}
    same, errs = True, {"droplets": [], "triple_point": []}
    for r in res_sat:
        cfg = by_rid[r.rid].cfg
        si, sj = cfg.grid.interior()
        same &= bool(np.array_equal(r.state[:, si, sj], poi_states[r.rid][:, si, sj]))
        kind = "triple_point" if cfg.name.startswith("triple") else "droplets"
        errs[kind].append(accuracy.compare(cfg, r.state, plain_states[r.rid]))
    keys = ("rho", "p", "alpha1", "centroid_shift_cells")
    cells = sum(r.cfg.grid.nx * r.cfg.grid.ny for r in reqs)
    return {
        "workload": f"{n} requests, seed {seed}: NX log-uniform in [128, 2048], NY = NX/2 "
                    f"(128..1024); droplets (MP5, t_end {serve.T_END_DROPLET}) and triple "
                    f"points (MP5, t_end {serve.T_END_TRIPLE}); initial states built on the "
                    f"device when a request starts (inside the timed runs)",
        "mean_cells": round(cells / n),
        "plain_fp64": {k: round(v, 4) for k, v in plain.items()},
        "batched_mixed_saturated": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in sat.items()},
        "batched_mixed_poisson": dict({k: (round(v, 4) if isinstance(v, float) else v) for k, v in poi.items()},
                                      rate_per_s=round(rate, 4), load=load),
        "speedup_cases_per_hour": round(sat["cases_per_hour"] / plain["cases_per_hour"], 3),
        "mix_invariant_bitwise": same,
        "max_rel_l1_vs_fp64": {kind: {k: float(max(e[k] for e in v)) for k in keys}
                               for kind, v in errs.items() if v},
    }


def _record(name: str) -> dict:
    """key=value lines of a committed measurement record under profiles/."""
    path = os.path.join(ROOT, "profiles", name)
    vals = {}
    if os.path.exists(path):
        with open(path) as fh:
            for line in fh:
                line = line.strip()
                if line and not line.startswith("#") and "=" in line:
                    k, v = line.split("=", 1)
                    vals[k.strip()] = v.strip()
    return vals


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE measurement (profiles/), or None when
    there is none for this workload."""
    vals = _record(f"{kernel}_traffic_{workload}.txt")
    if vals.get("workload") != workload or vals.get("kernel") != kernel:
        return None
    return int(vals["traffic_bytes"])


def copy_rate_gbs(mib: int = 1024, reps: int = 5) -> float:
    """Device-to-device copy rate of a MIB-MiB buffer (read + write bytes per
    second, GB/s), timed with events over REPS copies after one warm-up: the
    HBM rate a plain streaming kernel reaches on this device."""
    import torch

    a = torch.ones(mib * (1 << 20) // 8, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2.0 * a.numel() * 8 * reps / (e0.elapsed_time(e1) / 1e3) / 1e9
    del a, b
    torch.cuda.empty_cache()
    return gbs


def limiter_rates(workload: str, kernel: str, launch_ms: float, copy_gbs: float) -> dict:
    """The limiter launch's HBM rates: algorithmic bytes and the committed
    rocprofv3 traffic per launch (profiles/), each over the live launch time,
    beside the published peak and the measured copy rate."""
    vals = _record(f"{kernel}_traffic_{workload}.txt")
    out = {"kernel": kernel, "launch_ms": round(launch_ms, 4)}
    if vals.get("workload") != workload or vals.get("kernel") != kernel:
        return dict(out, traffic=None)
    t = launch_ms / 1e3
    alg, traffic = int(vals["alg_bytes"]), int(vals["traffic_bytes"])
    return dict(out, alg_bytes=alg, traffic=traffic,
                alg_gbs=round(alg / t / 1e9, 1), traffic_gbs=round(traffic / t / 1e9, 1),
                traffic_frac_of_peak=round(traffic / t / 1e9 / HBM_PEAK_GBS, 4),
                traffic_frac_of_copy=round(traffic / t / 1e9 / copy_gbs, 4))


def measured_step_work(workload: str):
    """(FP64 FLOP, VALU wave instructions) of one FP64 step of `workload` from
    the committed rocprofv3 instruction counts, or None."""
    vals = _record(f"step_fp64_work_{workload}.txt")
    if vals.get("workload") != workload:
        return None
    return int(vals["fp64_flop_per_step"]), int(vals["valu_wave_instr_per_step"])


def kernel_valu(workload: str, kernel: str, launch_ms: float) -> dict:
    """The kernel's VALU issue and FP64 FLOP rates at its live launch time, from
    the committed per-launch instruction counts (the kernel is VALU-bound; the
    contract's roofline object is HBM-framed)."""
    vals = _record(f"{kernel}_traffic_{workload}.txt")
    if vals.get("workload") != workload or "valu_wave_instr" not in vals:
        return {}
    t = launch_ms / 1e3
    valu, flop = int(vals["valu_wave_instr"]), int(vals["fp64_flop"])
    return {"valu_issue_frac": round(valu * 4 / (VALU_SIMDS * CLOCK_HZ * t), 4),
            "fp64_tflops": round(flop / t / 1e12, 3),
            "fp64_frac": round(flop / t / (FP64_PEAK_TFLOPS * 1e12), 4)}


def valu_fracs(workload: str, ms_per_step: float):
    """FP64 FLOP rate and VALU issue rate of the measured step against the
    chip's peaks (78.6 TFLOP/s FP64; 1024 SIMDs x 2.4 GHz / 4 cycles per wave
    instruction), from the committed per-step instruction counts."""
    work = measured_step_work(workload)
    if work is None:
        return None
    flop, valu = work
    t = ms_per_step / 1e3
    return {"fp64_tflops": round(flop / t / 1e12, 3),
            "fp64_frac": round(flop / t / (FP64_PEAK_TFLOPS * 1e12), 4),
            "valu_issue_frac": round(valu * 4 / (VALU_SIMDS * CLOCK_HZ * t), 4),
            "source": f"profiles/step_fp64_work_{workload}.txt"}


def late_window(s, done: int, start: int, steps: int) -> dict:
    """The headline solver, already `done` steps in, advanced untimed to step
    `start` (the shock in the water cylinder: it reaches the cylinder's front
    near step 680 and leaves its back near step 3500 at 4096x2048), then
    `steps` steps timed the same way as the headline window."""
    import ctypes as C

    import torch

    from wamp import _lib

    for _ in range(max(0, start - done)):
        s.step_device()
    torch.cuda.synchronize()
    _lib.check(s.L.wamp_timing_enable(s.ctx, steps), "wamp_timing_enable")
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        s.step_device()
    e1.record()
    torch.cuda.synchronize()
    t_ms = e0.elapsed_time(e1)
    ms = (C.c_double * 2)()
    nl = (C.c_int64 * 2)()
    _lib.check(s.L.wamp_timing_read(s.ctx, ms, nl), "wamp_timing_read")
    _lib.check(s.L.wamp_timing_enable(s.ctx, 0), "wamp_timing_enable")
    g = s.cfg.grid
    return {"first_step": max(start, done), "steps": steps,
            "value": round(g.nx * g.ny * steps / (t_ms / 1e3) / 1e9, 6),
            "ms_per_step": round(t_ms / steps, 4),
            "kernels_ms": {"k_faces": round(ms[0] / max(nl[0], 1), 4),
                           "k_limit_update": round(ms[1] / max(nl[1], 1), 4)},
            "state_finite": bool(torch.isfinite(s.U).all().item())}


def _progress(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=4096)
    ap.add_argument("--ny", type=int, default=2048)
    ap.add_argument("--cpu-steps", type=int, default=5,
                    help="steps of the CPU oracle sample (0 = skip the cpu_baseline leg)")
    ap.add_argument("--cpu-device-steps", type=int, default=2,
                    help="steps of the product CPU device (--cpu) on the same case (0 = skip)")
    ap.add_argument("--mixed-steps", type=int, default=50,
                    help="steps of the FP32-mixed leg on the same case (0 = skip)")
    ap.add_argument("--batch-cases", type=int, default=64,
                    help="cases of the batched leg (0 = skip)")
    ap.add_argument("--batch-steps", type=int, default=500, help="steps per case in the batched leg")
    ap.add_argument("--serve-requests", type=int, default=200,
                    help="requests of the serving leg (0 = skip)")
    ap.add_argument("--threed-n", type=int, default=160,
                    help="the 3D leg's shock-sphere case has 2n x n x n cells")
    ap.add_argument("--threed-steps", type=int, default=20,
                    help="timed steps of the 3D leg (0 = skip)")
    ap.add_argument("--threed-batch-cases", type=int, default=32,
                    help="uneven 3D cases in the 3D batched leg (0: skip)")
    ap.add_argument("--threed-batch-steps", type=int, default=100)
    ap.add_argument("--late-start", type=int, default=3500,
                    help="also time --steps steps of the headline case from this step on, after "
                         "the shock has crossed the cylinder (0 = skip)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="before the run, copy a MIB-MiB device buffer once (a plain copy kernel "
                         "of known size, to scale hardware traffic counters)")
    args = ap.parse_args(argv)

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one case per GPU; "
              "a single case would slab-decompose with halo exchange, not implemented). "
              "Run with --gpus 1.", file=sys.stderr)
        return 2

    import torch

    from wamp import cases
    from wamp.config import SCHEME_M10
    from wamp import _lib
    from wamp.solver import Solver

    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        return 2
    if args.calibrate_copy > 0:
        n = args.calibrate_copy * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    copy_gbs = copy_rate_gbs()
    cfg, U0 = cases.mach10_cylinder(args.nx, args.ny, scheme=SCHEME_M10)
    g = cfg.grid
    s = Solver(cfg, device="cuda:0")
    s.set_state(U0)
    s.prepare()
    s.compute_dt()
    for _ in range(args.warmup):
        s.step_device()
    torch.cuda.synchronize()

    K = args.steps
    _lib.check(s.L.wamp_timing_enable(s.ctx, K), "wamp_timing_enable")
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(K):
        s.step_device()
    e1.record(stream)
    torch.cuda.synchronize()
    t_ms = e0.elapsed_time(e1)
    import ctypes as C

    ms = (C.c_double * 2)()
    nl = (C.c_int64 * 2)()
    _lib.check(s.L.wamp_timing_read(s.ctx, ms, nl), "wamp_timing_read")
    _lib.check(s.L.wamp_timing_enable(s.ctx, 0), "wamp_timing_enable")
    dt_now = float(s.dt_buf[0].item())
    finite = bool(torch.isfinite(s.U).all().item())
    cells = g.nx * g.ny
    value = cells * K / (t_ms / 1e3) / 1e9
    faces_ms = ms[0] / max(nl[0], 1)
    lu_ms = ms[1] / max(nl[1], 1)
    # algorithmic bytes of one face-kernel launch: one stage's share of the
    # step's 384 B per cell-update, over all NX*NY cells
    alg_bytes = ALG_BYTES_PER_CELL_UPDATE / 3.0 * cells
    achieved = alg_bytes / (faces_ms / 1e3) / 1e9
    step_gbs = ALG_BYTES_PER_CELL_UPDATE * cells * K / (t_ms / 1e3) / 1e9
    kvalu = kernel_valu(f"mach10_cylinder_{g.nx}x{g.ny}", "k_faces", faces_ms)

    out = {
        "metric": "Gcell-updates/s",
        "value": round(value, 6),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(t_ms / K, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"mach10_cylinder_{g.nx}x{g.ny}",
            "grid": [g.nx, g.ny],
            "scheme": "M10 (MUSCL + THINC, Wong face+flux limiters)",
            "precision": "fp64-plain",
            "cfl": cfg.scheme.cfl,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_faces",
            "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": measured_traffic(f"mach10_cylinder_{g.nx}x{g.ny}", "k_faces"),
            "launch_ms": round(faces_ms, 4),
            "launches": int(nl[0]),
            **kvalu,
            # the bound that binds: the face kernel is limited by FP64 VALU
            # issue, not HBM (the HBM fraction above is the algorithmic-bytes
            # framing); valu_issue_frac = issued VALU wave instructions x 4
            # cycles / (1024 SIMDs x 2.4 GHz x launch time)
            "binding": {"bound": "valu_issue", "frac": kvalu.get("valu_issue_frac"),
                        "fp64_flop_frac": kvalu.get("fp64_frac")},
        },
        "kernels_ms": {"k_faces": round(faces_ms, 4), "k_limit_update": round(lu_ms, 4)},
        # the FP64 step's limiter at this size: the march that forms the
        # first-order fluxes itself (the face kernel writes only the 14
        # high-order planes)
        "limiter": limiter_rates(f"mach10_cylinder_{g.nx}x{g.ny}", "k_lu_lo", lu_ms, copy_gbs),
        "hbm_copy_gbs": round(copy_gbs, 1),
        "step_hbm_frac": round(step_gbs / HBM_PEAK_GBS, 5),
        "step_valu_frac": valu_fracs(f"mach10_cylinder_{g.nx}x{g.ny}", t_ms / K),
        "state_finite": finite,
        "dt_last": dt_now,
    }
    if args.late_start > 0:
        out["late_window"] = late_window(s, args.warmup + K, args.late_start, K)
        _progress(f"late window from step {out['late_window']['first_step']}: "
                  f"{out['late_window']['value']:.4f} Gcell-updates/s")
    if args.mixed_steps > 0:
        out["mixed"] = mixed_leg(cfg, U0, args.warmup, args.mixed_steps)
    s.close()
    _progress(f"headline {out['value']:.4f} Gcell-updates/s")
    if args.threed_steps > 0:
        out["threed"] = threed_leg(args.threed_n, 3, args.threed_steps)
        _progress(f"3D {out['threed']['value']:.4f} Gcell-updates/s")
    if args.threed_batch_cases > 0:
        out["threed_batched"] = threed_batched_leg(args.threed_batch_cases, args.threed_batch_steps)
        _progress(f"3D batched speedup {out['threed_batched']['speedup']}")
    if args.batch_cases > 0:
        out["batched"] = batched_leg(args.batch_cases, args.batch_steps)
        _progress(f"batched speedup {out['batched']['speedup']}")
    if args.serve_requests > 0:
        out["serving"] = serving_leg(args.serve_requests)
    if args.cpu_steps > 0:
        out["cpu_baseline"] = cpu_baseline(cfg, U0, args.cpu_steps)
    else:
        out["cpu_baseline"] = None
    if args.cpu_device_steps > 0:
        out["cpu_device"] = cpu_device_leg(cfg, U0, args.cpu_device_steps)
        _progress(f"CPU device {out['cpu_device']['value']:.5f} Gcell-updates/s")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
