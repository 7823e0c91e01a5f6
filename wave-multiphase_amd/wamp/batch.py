"""Batched solves: many independent cases of uneven size advanced together.

One persistent launch per kernel kind per RK stage walks a work list made of
every case's face tiles (face kernel) and limiter rows (update kernel); the
per-case dt, time and end time live on the device and a whole step (or
several) is captured in a hipGraph.  `autotune` picks the tile shape, the
persistent blocks per CU and the limiter rows per wave by timing the batch.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from . import _lib
from .config import CaseConfig


def _torch():
    import torch

    return torch


@dataclass
class Tuning:
    tile: int = 0          # 0: 8x32 face tiles, 1: 16x16
    faces_bpcu: int = 3    # persistent face blocks per CU
    lu_bpcu: int = 0       # persistent limiter blocks (4 waves) per CU; 0: one wave per item
    lu_seg: int = 16       # limiter rows per wave; 0: chosen at each commit for the resident cases

    def as_tuple(self):
        return (self.tile, self.faces_bpcu, self.lu_bpcu, self.lu_seg)


TUNING_SPACE = [Tuning(t, f, l, s) for t in (0, 1) for f in (2, 3) for l in (0, 2) for s in (0, 26, 48)]


class Batch:
    """Slots of resident cases sharing one scheme and precision."""

    def __init__(self, scheme, precision: int, max_cases: int, device: str = "cuda"):
        torch = _torch()
        if not torch.cuda.is_available():
            raise _lib.WampError("batches need a ROCm GPU")
        self.L = _lib.load()
        self.device = torch.device(device)
        self.scheme = scheme
        self.precision = precision
        self.max_cases = max_cases
        sch = _lib.WampScheme(scheme.base, scheme.shear, scheme.vf_thinc, scheme.face_pp,
                              1 if scheme.repair else 0, scheme.cfl)
        h = C.c_void_p()
        _lib.check(self.L.wamp_batch_create(C.byref(sch), int(precision), int(max_cases), C.byref(h)),
                   "wamp_batch_create")
        self.h = h
        self.cfgs: dict[int, CaseConfig] = {}
        self.states: dict[int, object] = {}
        self.t0: dict[int, float] = {}
        self.tuning = Tuning()
        self.graph_steps = 0

    @property
    def stream(self) -> int:
        return _torch().cuda.current_stream(self.device).cuda_stream

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.wamp_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- slots ---------------------------------------------------------------
    def set_case(self, slot: int, cfg: CaseConfig, U0: np.ndarray, t0: float = 0.0,
                 t_end: float = 0.0):
        torch = _torch()
        g = cfg.grid
        if cfg.precision != self.precision or cfg.scheme != self.scheme:
            raise ValueError("a batch holds one scheme and one precision")
        if isinstance(U0, torch.Tensor) and U0.device == self.device and \
                tuple(U0.shape) == (6, g.nxg, g.nyg) and U0.dtype == torch.float64 and U0.is_contiguous():
            U = U0  # a device state built for this slot (cases.*_device): kept, not copied
        else:
            U = torch.zeros((6, g.nxg, g.nyg), dtype=torch.float64, device=self.device)
            if isinstance(U0, torch.Tensor):
                U.copy_(U0[:, :, : g.nyg])
            else:
                U.copy_(torch.from_numpy(np.ascontiguousarray(U0, dtype=np.float64)))
        grid = _lib.WampGrid(g.nx, g.ny, g.g, g.nyg, g.dx, g.dy, (C.c_int32 * 4)(*g.bc))
        eos = _lib.WampEos(*cfg.eos.as_tuple())
        ptrs = _lib.ptrs6([U[m].data_ptr() for m in range(6)])
        _lib.check(self.L.wamp_batch_set_case(self.h, int(slot), C.byref(grid), C.byref(eos), ptrs,
                                              float(t0), float(t_end)), "wamp_batch_set_case")
        self.cfgs[slot] = cfg
        self.states[slot] = U
        self.t0[slot] = float(t0)

    def clear_case(self, slot: int):
        _lib.check(self.L.wamp_batch_clear_case(self.h, int(slot)), "wamp_batch_clear_case")
        self.cfgs.pop(slot, None)
        self.states.pop(slot, None)
        self.t0.pop(slot, None)

    def state(self, slot: int) -> np.ndarray:
        return self.states[slot].cpu().numpy().copy()

    # -- execution -----------------------------------------------------------
    def set_tuning(self, tuning: Tuning):
        _lib.check(self.L.wamp_batch_set_tuning(self.h, *tuning.as_tuple()), "wamp_batch_set_tuning")
        self.tuning = tuning

    def commit(self, graph_steps: int = 0):
        _lib.check(self.L.wamp_batch_commit(self.h, self.stream), "wamp_batch_commit")
        self.graph_steps = 0
        if graph_steps > 0:
            self.capture(graph_steps)

    def capture(self, steps: int):
        torch = _torch()
        torch.cuda.synchronize(self.device)
        _lib.check(self.L.wamp_batch_graph(self.h, int(steps)), "wamp_batch_graph")
        self.graph_steps = steps

    def prepare(self, slots=None):
        if slots is None:
            _lib.check(self.L.wamp_batch_prepare(self.h, None, -1, self.stream), "wamp_batch_prepare")
        else:
            arr = (C.c_int32 * len(slots))(*slots)
            _lib.check(self.L.wamp_batch_prepare(self.h, arr, len(slots), self.stream), "wamp_batch_prepare")

    def set_done(self, slot: int, done: bool = True):
        _lib.check(self.L.wamp_batch_set_done(self.h, int(slot), 1 if done else 0, self.stream),
                   "wamp_batch_set_done")

    def step(self) -> int:
        rc = self.L.wamp_batch_step(self.h, self.stream)
        if rc <= 0:
            _lib.check(rc if rc != 0 else -1, "wamp_batch_step")
        return rc

    def status(self):
        n = self.max_cases
        t = (C.c_double * n)()
        dt = (C.c_double * n)()
        done = (C.c_int32 * n)()
        _lib.check(self.L.wamp_batch_status(self.h, t, dt, done, self.stream), "wamp_batch_status")
        return np.array(t), np.array(dt), np.array(done, dtype=bool)

    def work(self):
        """(face tiles, limiter wave items, limiter rows per wave) of the committed table."""
        return tuple(int(self.L.wamp_batch_work(self.h, w)) for w in range(3))

    # -- tuning ----------------------------------------------------------------
    def autotune(self, steps: int = 3, space=None, log=None) -> Tuning:
        """Time `steps` steps per candidate on a snapshot of the states; keep the fastest.

        The work decomposition does not change any result, so the snapshot is
        only restored to leave the cases where they were.  The restore also
        resets each case's time to its start time (prepare), so autotune runs
        before the batch steps, and refuses otherwise."""
        torch = _torch()
        t_before, _, _ = self.status()
        moved = [k for k in self.states if t_before[k] != self.t0[k]]
        if moved:
            raise RuntimeError(f"autotune runs before the batch steps (slots {moved} have advanced)")
        snap = {k: v.clone() for k, v in self.states.items()}
        best, best_ms = None, float("inf")
        for cand in (space or TUNING_SPACE):
            self.set_tuning(cand)
            self.commit()
            self.step()  # warm-up (first launch of this shape)
            torch.cuda.synchronize(self.device)
            t0 = time.perf_counter()
            for _ in range(steps):
                self.step()
            torch.cuda.synchronize(self.device)
            ms = (time.perf_counter() - t0) * 1e3 / steps
            if log:
                log(f"  tuning {cand.as_tuple()}: {ms:.3f} ms/step")
            if ms < best_ms:
                best, best_ms = cand, ms
            for k, v in snap.items():
                self.states[k].copy_(v)
            self.prepare(list(self.states))  # restores t, dt and done from t0
        self.set_tuning(best)
        self.commit()
        self.prepare(list(self.states))
        self.best_ms = best_ms
        return best
