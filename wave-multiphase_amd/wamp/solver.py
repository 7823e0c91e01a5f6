"""Host side of the HIP solver step: the MI355X counterpart of main() (SC:1828-2082).

Device buffers are PyTorch-ROCm tensors; every kernel runs through
libwamp.so's C-ABI on the current torch stream.  There is no host or
PyTorch fallback for any part of the step.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .config import CaseConfig


def _torch():
    import torch

    return torch


class Solver:
    """One resident case: conserved state (6 FP64 planes), dt on the device, a libwamp context."""

    def __init__(self, cfg: CaseConfig, device: str = "cuda", ld: int | None = None):
        torch = _torch()
        if not torch.cuda.is_available():
            raise _lib.WampError("the HIP solver needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.L = _lib.load()
        self.cfg = cfg
        g = cfg.grid
        self.ld = int(ld) if ld else g.nyg
        if self.ld < g.nyg:
            raise ValueError("ld must be >= ny + 2g")
        self.device = torch.device(device)
        self.U = torch.zeros((6, g.nxg, self.ld), dtype=torch.float64, device=self.device)
        self.dt_buf = torch.zeros(2, dtype=torch.float64, device=self.device)
        self._u = _lib.ptrs6([self.U[m].data_ptr() for m in range(6)])
        grid = _lib.WampGrid(g.nx, g.ny, g.g, self.ld, g.dx, g.dy, (C.c_int32 * 4)(*g.bc))
        eos = _lib.WampEos(*cfg.eos.as_tuple())
        s = cfg.scheme
        sch = _lib.WampScheme(s.base, s.shear, s.vf_thinc, s.face_pp, 1 if s.repair else 0, s.cfl)
        ctx = C.c_void_p()
        _lib.check(self.L.wamp_ctx_create(C.byref(grid), C.byref(eos), C.byref(sch),
                                          int(cfg.precision), C.byref(ctx)), "wamp_ctx_create")
        self.ctx = ctx
        self.time = 0.0
        self.nstep = 0

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "ctx", None) and self.ctx.value:
            _lib.check(self.L.wamp_ctx_destroy(self.ctx), "wamp_ctx_destroy")
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return _torch().cuda.current_stream(self.device).cuda_stream

    @property
    def dt_ptr(self) -> int:
        return self.dt_buf.data_ptr()

    def scratch_bytes(self) -> int:
        return int(self.L.wamp_ctx_scratch_bytes(self.ctx))

    # -- state --------------------------------------------------------------
    def set_state(self, U):
        """Set the (6, NXG, NYG) conserved state: a host FP64 array (uploaded) or a
        device tensor (copied on the device, e.g. from cases.*_device)."""
        g = self.cfg.grid
        if tuple(U.shape[:2]) != (6, g.nxg) or U.shape[2] < g.nyg:
            raise ValueError(f"state shape {tuple(U.shape)} != {(6, g.nxg, g.nyg)}")
        torch = _torch()
        if isinstance(U, torch.Tensor):
            self.U[:, :, : g.nyg].copy_(U[:, :, : g.nyg])
        else:
            self.U[:, :, : g.nyg].copy_(torch.from_numpy(np.ascontiguousarray(U, dtype=np.float64)))

    def state(self) -> np.ndarray:
        g = self.cfg.grid
        return self.U[:, :, : g.nyg].cpu().numpy().copy()

    def prims(self) -> np.ndarray:
        """(3, NXG, NYG) u, v, p via cons_to_prim (SC:840-864) after a boundary fill."""
        torch = _torch()
        g = self.cfg.grid
        P = torch.zeros((3, g.nxg, self.ld), dtype=torch.float64, device=self.device)
        p3 = _lib.ptrs3([P[m].data_ptr() for m in range(3)])
        _lib.check(self.L.wamp_apply_bc(self.ctx, self._u, None, self.stream), "wamp_apply_bc")
        _lib.check(self.L.wamp_cons_to_prim(self.ctx, self._u, p3, self.stream), "wamp_cons_to_prim")
        return P[:, :, : g.nyg].cpu().numpy().copy()

    # -- the step's closures (SC:1894-1995) ----------------------------------
    def prepare(self):
        """Initial sequence SC:1998-2001 (repair + alpha clamp; ghosts via apply_bc)."""
        _lib.check(self.L.wamp_prepare(self.ctx, self._u, self.stream), "wamp_prepare")
        _lib.check(self.L.wamp_apply_bc(self.ctx, self._u, None, self.stream), "wamp_apply_bc")

    def apply_bc(self):
        _lib.check(self.L.wamp_apply_bc(self.ctx, self._u, None, self.stream), "wamp_apply_bc")

    def compute_dt(self):
        """dt = CFL * min(...) into the device dt slot (get_dt SC:1986-1995, no host sync)."""
        _lib.check(self.L.wamp_get_dt(self.ctx, self._u, self.dt_ptr, self.stream), "wamp_get_dt")

    def get_dt(self) -> float:
        self.compute_dt()
        return float(self.dt_buf[0].item())

    def set_dt(self, dt: float):
        self.dt_buf[0].fill_(float(dt))

    def step_device(self):
        """One SSP-RK3 step with the device-resident dt; the next dt replaces it."""
        _lib.check(self.L.wamp_step(self.ctx, self._u, self.dt_ptr, self.dt_ptr, self.stream),
                   "wamp_step")
        self.nstep += 1

    def fluxes(self):
        """The 28 face arrays of compute_fluxes + compute_low_fluxes (SC:1919-1947)."""
        torch = _torch()
        g = self.cfg.grid
        F = torch.zeros((28, g.nxg, self.ld), dtype=torch.float64, device=self.device)
        _lib.check(self.L.wamp_compute_fluxes(self.ctx, self._u,
                                              _lib.ptrs28([F[m].data_ptr() for m in range(28)]),
                                              self.stream), "wamp_compute_fluxes")
        return F[:, :, : g.nyg].cpu().numpy().copy()

    def run(self, nsteps: int) -> np.ndarray:
        """n steps with dt from the state before each step (SC:2015, 2059); returns the dts."""
        torch = _torch()
        dts = torch.zeros(nsteps, dtype=torch.float64, device=self.device)
        self.compute_dt()
        for k in range(nsteps):
            dts[k].copy_(self.dt_buf[0])
            self.step_device()
        return dts.cpu().numpy()

    def step_dt(self, dt: float) -> float:
        """One step with time step dt; returns the next dt (cfl * min of the new state)."""
        self.set_dt(dt)
        self.step_device()
        return float(self.dt_buf[0].item())

    def advance_to(self, t_end: float, save_times=(), on_save=None, log_every: int = 200,
                   log=print, max_steps: int | None = None):
        """Time loop of main() (SC:2010-2073): dt clipped to save times and t_end."""
        return advance(self, t_end, save_times, on_save, log_every, log, max_steps)


def advance(s, t_end: float, save_times=(), on_save=None, log_every: int = 200, log=print,
            max_steps: int | None = None):
    """Time loop of main() (SC:2010-2073) over a solver `s` with `time`, `nstep`,
    get_dt() and step_dt(dt) -> next dt: dt clipped to the save times and t_end,
    on_save(s) at each save time."""
    import time as _time

    save_flags = [s.time >= ts - 1.0e-14 for ts in save_times]
    dt = s.get_dt()
    t0 = _time.time()
    while s.time < t_end - 1.0e-14:
        if max_steps is not None and s.nstep >= max_steps:
            break
        for k, ts in enumerate(save_times):
            if not save_flags[k] and s.time + dt > ts:
                dt = ts - s.time
        if s.time + dt > t_end:
            dt = t_end - s.time
        step_dt = dt
        dt = s.step_dt(step_dt)
        s.time += step_dt
        if log_every and s.nstep % log_every == 0 and log:
            log(f"  step {s.nstep:5d}  t={s.time:.4e}  dt={dt:.3e}  "
                f"wall={_time.time() - t0:.1f}s")
        for k, ts in enumerate(save_times):
            if not save_flags[k] and abs(s.time - ts) < 0.5 * dt:
                if on_save:
                    on_save(s)
                save_flags[k] = True
    return s.time
