"""Three-dimensional extension: grid, cases and the resident solver over
libwamp's wamp3_* entry points (include/wamp.h).

The reference's scripts are 2D (SC, M10); this is the same step with a
z-sweep and the 7-variable state [a1 rho1, a2 rho2, rho u, rho v, rho w, E,
a1], laid out (7, nx+2g, ny+2g, ldz) with z contiguous.  `from_2d_xy` and
`from_2d_xz` embed a 2D state as a z- or y-invariant 3D one; the 3D step
of such a state equals the 2D step of the slice exactly (for the same dt),
which is how the 3D kernels are pinned to the 2D path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .cases import ALPHA_CLIP, WATER_ND, rankine_hugoniot, _smeared
from .config import (BC_SLIP_WALL, BC_ZERO_GRADIENT, EOS, G_DEFAULT, SCHEME_M10, CaseConfig,
                     Grid, Scheme)

# conserved component of the 3D state holding each 2D one (ar1, ar2, ru, rv, E, a1)
XY_MAP = (0, 1, 2, 3, 5, 6)   # z-invariant embedding: 2D (x, y) -> 3D (x, y)
XZ_MAP = (0, 1, 2, 4, 5, 6)   # y-invariant embedding: 2D (x, y) -> 3D (x, z), 2D v -> w


@dataclass(frozen=True)
class Grid3:
    """Uniform cell-centred 3D grid with g ghost layers on every side."""

    nx: int
    ny: int
    nz: int
    xmin: float = 0.0
    xmax: float = 1.0
    ymin: float = 0.0
    ymax: float = 1.0
    zmin: float = 0.0
    zmax: float = 1.0
    g: int = G_DEFAULT
    # x-lo, x-hi, y-lo, y-hi, z-lo, z-hi
    bc: tuple = (BC_ZERO_GRADIENT, BC_ZERO_GRADIENT, BC_SLIP_WALL, BC_SLIP_WALL,
                 BC_SLIP_WALL, BC_SLIP_WALL)

    @property
    def nxg(self) -> int:
        return self.nx + 2 * self.g

    @property
    def nyg(self) -> int:
        return self.ny + 2 * self.g

    @property
    def nzg(self) -> int:
        return self.nz + 2 * self.g

    @property
    def dx(self) -> float:
        return (self.xmax - self.xmin) / self.nx

    @property
    def dy(self) -> float:
        return (self.ymax - self.ymin) / self.ny

    @property
    def dz(self) -> float:
        return (self.zmax - self.zmin) / self.nz

    def centres(self, axis: int):
        n, lo, d = [(self.nxg, self.xmin, self.dx), (self.nyg, self.ymin, self.dy),
                    (self.nzg, self.zmin, self.dz)][axis]
        return lo + (np.arange(n, dtype=np.float64) - self.g + 0.5) * d

    def interior(self):
        return (slice(self.g, self.g + self.nx), slice(self.g, self.g + self.ny),
                slice(self.g, self.g + self.nz))


@dataclass
class Case3Config:
    grid: Grid3
    eos: EOS = field(default_factory=lambda: WATER_ND)
    scheme: Scheme = SCHEME_M10
    name: str = "case3"
    precision: int = 0  # 0 FP64, 1 FP32-mixed face arithmetic


def conserved3(eos: EOS, a1, rho1, rho2, u, v, w, p):
    """[ar1, ar2, ru, rv, rw, E, a1], E as prim_to_energy (SC:105-110) with w."""
    ar1 = a1 * rho1
    ar2 = (1.0 - a1) * rho2
    rho = ar1 + ar2
    a2 = 1.0 - a1
    gb = a1 / (eos.gamma1 - 1.0) + a2 / (eos.gamma2 - 1.0)
    pb = a1 * eos.gamma1 * eos.pinf1 / (eos.gamma1 - 1.0) + a2 * eos.gamma2 * eos.pinf2 / (eos.gamma2 - 1.0)
    E = gb * p + pb + 0.5 * rho * ((u * u + v * v) + w * w)
    return np.stack([ar1, ar2, rho * u, rho * v, rho * w, E, a1]).astype(np.float64)


def shock_sphere(n: int = 128, scheme: Scheme = SCHEME_M10, mach: float = 2.4):
    """Shock at x = 0.4 moving +x through air into a water sphere r = 0.15 at
    (0.75, 0.5, 0.5); domain [0, 2] x [0, 1] x [0, 1] on 2n x n x n cells."""
    grid = Grid3(2 * n, n, n, 0.0, 2.0, 0.0, 1.0, 0.0, 1.0)
    cfg = Case3Config(grid, WATER_ND, scheme, name=f"shock_sphere_{2 * n}x{n}x{n}")
    x = grid.centres(0)[:, None, None]
    y = grid.centres(1)[None, :, None]
    z = grid.centres(2)[None, None, :]
    d = np.sqrt((x - 0.75) ** 2 + ((y - 0.5) ** 2 + (z - 0.5) ** 2))  # exactly y <-> z symmetric
    a1 = np.clip(_smeared(d, 0.15, 1.5 * grid.dx), ALPHA_CLIP, 1.0 - ALPHA_CLIP)
    rho_s, p_s, u_s = rankine_hugoniot(mach, WATER_ND.gamma2, 1.0, 1.0)
    post = np.broadcast_to(x < 0.4, a1.shape)
    rho2 = np.where(post, rho_s, 1.0)
    p = np.where(post, p_s, 1.0)
    u = np.where(post, u_s, 0.0)
    zero = np.zeros_like(u)
    U = conserved3(WATER_ND, a1, np.full_like(a1, 1000.0), rho2, u, zero, zero, p)
    out = np.zeros_like(U)
    si, sj, sk = grid.interior()
    out[:, si, sj, sk] = U[:, si, sj, sk]
    return cfg, out


def from_2d_xy(cfg2: CaseConfig, U2: np.ndarray, nz: int = 4, bc_z=(BC_SLIP_WALL, BC_SLIP_WALL)):
    """The 2D state as a z-invariant 3D state with w = 0 (dz = dx)."""
    g2: Grid = cfg2.grid
    dz = g2.dx
    grid = Grid3(g2.nx, g2.ny, nz, g2.xmin, g2.xmax, g2.ymin, g2.ymax, 0.0, nz * dz, g2.g,
                 tuple(g2.bc) + tuple(bc_z))
    U = np.zeros((7, g2.nxg, g2.nyg, grid.nzg))
    for m2, m3 in enumerate(XY_MAP):
        U[m3] = np.asarray(U2[m2][:, : g2.nyg])[:, :, None]
    return Case3Config(grid, cfg2.eos, cfg2.scheme, name=cfg2.name + "_xy"), U


def from_2d_xz(cfg2: CaseConfig, U2: np.ndarray, ny: int = 4, bc_y=(BC_SLIP_WALL, BC_SLIP_WALL)):
    """The 2D state as a y-invariant 3D state with v = 0: 2D y becomes z and
    2D v becomes w (dy = dx)."""
    g2: Grid = cfg2.grid
    dy = g2.dx
    grid = Grid3(g2.nx, ny, g2.ny, g2.xmin, g2.xmax, 0.0, ny * dy, g2.ymin, g2.ymax, g2.g,
                 (g2.bc[0], g2.bc[1]) + tuple(bc_y) + (g2.bc[2], g2.bc[3]))
    U = np.zeros((7, g2.nxg, grid.nyg, g2.nyg))
    for m2, m3 in enumerate(XZ_MAP):
        U[m3] = np.asarray(U2[m2][:, : g2.nyg])[:, None, :]
    return Case3Config(grid, cfg2.eos, cfg2.scheme, name=cfg2.name + "_xz"), U


def pressure3(eos: EOS, U):
    """Pressure of a (7, ...) state (cons_to_prim SC:856-860 with w), numpy."""
    al = U[6]
    rho = U[0] + U[1]
    u, v, w = U[2] / rho, U[3] / rho, U[4] / rho
    a2 = 1.0 - al
    gb = al / (eos.gamma1 - 1.0) + a2 / (eos.gamma2 - 1.0)
    pb = al * eos.gamma1 * eos.pinf1 / (eos.gamma1 - 1.0) + a2 * eos.gamma2 * eos.pinf2 / (eos.gamma2 - 1.0)
    return (U[5] - pb - 0.5 * rho * ((u * u + v * v) + w * w)) / gb


def compare3(cfg: Case3Config, U, Uref) -> dict:
    """Relative L1 errors of density, pressure and alpha1 over the interior."""
    from .accuracy import rel_l1

    si, sj, sk = cfg.grid.interior()
    I, R = U[:, si, sj, sk], Uref[:, si, sj, sk]
    return {"rho": rel_l1(I[0] + I[1], R[0] + R[1]),
            "p": rel_l1(pressure3(cfg.eos, I), pressure3(cfg.eos, R)),
            "alpha1": rel_l1(I[6], R[6]),
            "finite": bool(np.isfinite(I).all())}


REQUIRED3 = ["ar1", "ar2", "u", "v", "w", "p", "alpha1"]


def save_npz3(path: str, grid: Grid3, U, t: float, eos: EOS) -> str:
    """Interior of a 3D state in the npz layout of the 2D writer (save_npz,
    SC:1712-1726) with z and w added: keys time, x, y, z, rho, ar1, ar2, u, v,
    w, p, alpha1, arrays of shape (nx, ny, nz)."""
    si, sj, sk = grid.interior()
    I = np.asarray(U)[:, si, sj, sk]
    rho = I[0] + I[1]
    np.savez_compressed(path, time=np.float64(t),
                        x=grid.centres(0)[si], y=grid.centres(1)[sj], z=grid.centres(2)[sk],
                        rho=rho, ar1=I[0], ar2=I[1], u=I[2] / rho, v=I[3] / rho, w=I[4] / rho,
                        p=pressure3(eos, I), alpha1=I[6])
    return path


def load_restart3(path: str, cfg: Case3Config):
    """(U (7, NXG, NYG, NZG) with zero ghosts, time) from a save_npz3 file, as
    load_restart (SC:1779-1821): required keys checked, alpha1 clipped to
    [1e-8, 1 - 1e-8], E rebuilt from p with the run's EOS."""
    g, e = cfg.grid, cfg.eos
    with np.load(path, allow_pickle=False) as data:
        for key in REQUIRED3:
            if key not in data:
                raise KeyError(f"Restart file {path} is missing '{key}'")
        f = {k: np.asarray(data[k], dtype=np.float64) for k in REQUIRED3}
        t = float(data["time"]) if "time" in data else 0.0
    if f["ar1"].shape != (g.nx, g.ny, g.nz):
        raise ValueError(f"Restart shape {f['ar1'].shape} does not match {(g.nx, g.ny, g.nz)}")
    al = np.clip(f["alpha1"], 1.0e-8, 1.0 - 1.0e-8)
    rho = f["ar1"] + f["ar2"]
    a2 = 1.0 - al
    gb = al / (e.gamma1 - 1.0) + a2 / (e.gamma2 - 1.0)
    pb = al * e.gamma1 * e.pinf1 / (e.gamma1 - 1.0) + a2 * e.gamma2 * e.pinf2 / (e.gamma2 - 1.0)
    u, v, w = f["u"], f["v"], f["w"]
    E = gb * f["p"] + pb + 0.5 * rho * ((u * u + v * v) + w * w)
    U = np.zeros((7, g.nxg, g.nyg, g.nzg), dtype=np.float64)
    si, sj, sk = g.interior()
    for m, a in enumerate((f["ar1"], f["ar2"], rho * u, rho * v, rho * w, E, al)):
        U[m][si, sj, sk] = a
    return U, t


def _torch():
    import torch

    return torch


class Solver3:
    """One resident 3D case: conserved state (7 FP64 planes), dt on the device, a
    libwamp 3D context.  Every call enqueues on the current torch stream."""

    def __init__(self, cfg: Case3Config, device: str = "cuda", ldz: int | None = None):
        torch = _torch()
        if not torch.cuda.is_available():
            raise _lib.WampError("the HIP solver needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.L = _lib.load()
        self.cfg = cfg
        g = cfg.grid
        self.ldz = int(ldz) if ldz else g.nzg
        if self.ldz < g.nzg:
            raise ValueError("ldz must be >= nz + 2g")
        self.device = torch.device(device)
        self.U = torch.zeros((7, g.nxg, g.nyg, self.ldz), dtype=torch.float64, device=self.device)
        self.dt_buf = torch.zeros(2, dtype=torch.float64, device=self.device)
        self._u = _lib.PtrArr7(*[self.U[m].data_ptr() for m in range(7)])
        grid = _lib.WampGrid3(g.nx, g.ny, g.nz, g.g, self.ldz, g.dx, g.dy, g.dz,
                              (C.c_int32 * 6)(*g.bc))
        eos = _lib.WampEos(*cfg.eos.as_tuple())
        s = cfg.scheme
        sch = _lib.WampScheme(s.base, s.shear, s.vf_thinc, s.face_pp, 1 if s.repair else 0, s.cfl)
        ctx = C.c_void_p()
        _lib.check(self.L.wamp3_ctx_create(C.byref(grid), C.byref(eos), C.byref(sch),
                                           int(cfg.precision), C.byref(ctx)), "wamp3_ctx_create")
        self.ctx = ctx
        self.nstep = 0

    def close(self):
        if getattr(self, "ctx", None) and self.ctx.value:
            _lib.check(self.L.wamp3_ctx_destroy(self.ctx), "wamp3_ctx_destroy")
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
        return int(self.L.wamp3_ctx_scratch_bytes(self.ctx))

    def set_state(self, U):
        g = self.cfg.grid
        if tuple(U.shape[:3]) != (7, g.nxg, g.nyg) or U.shape[3] < g.nzg:
            raise ValueError(f"state shape {tuple(U.shape)} != {(7, g.nxg, g.nyg, g.nzg)}")
        torch = _torch()
        src = U if isinstance(U, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(U, dtype=np.float64))
        self.U[..., : g.nzg].copy_(src[..., : g.nzg])

    def state(self) -> np.ndarray:
        return self.U[..., : self.cfg.grid.nzg].cpu().numpy().copy()

    def prepare(self):
        _lib.check(self.L.wamp3_prepare(self.ctx, self._u, self.stream), "wamp3_prepare")

    def compute_dt(self):
        _lib.check(self.L.wamp3_get_dt(self.ctx, self._u, self.dt_ptr, self.stream), "wamp3_get_dt")

    def get_dt(self) -> float:
        self.compute_dt()
        return float(self.dt_buf[0].item())

    def set_dt(self, dt: float):
        self.dt_buf[0].fill_(float(dt))

    def step_device(self):
        """One SSP-RK3 step with the device-resident dt; the next dt replaces it."""
        _lib.check(self.L.wamp3_step(self.ctx, self._u, self.dt_ptr, self.dt_ptr, self.stream),
                   "wamp3_step")
        self.nstep += 1

    def step_with(self, dt: float):
        """One step with the given dt (the CFL dt of the result is discarded)."""
        self.set_dt(dt)
        _lib.check(self.L.wamp3_step(self.ctx, self._u, self.dt_ptr, self.dt_buf[1:].data_ptr(),
                                     self.stream), "wamp3_step")
        self.nstep += 1

    def step_with_device_dt(self):
        """One step with the dt already in the device slot, which it keeps (the
        CFL dt of the result goes to the second slot)."""
        _lib.check(self.L.wamp3_step(self.ctx, self._u, self.dt_ptr, self.dt_buf[1:].data_ptr(),
                                     self.stream), "wamp3_step")
        self.nstep += 1

    def fluxes(self):
        """The 48 face arrays of one stage's flux pass on the current state
        (x, y, z: 8 high-order fluxes with u*, then the 8 first-order ones)."""
        torch = _torch()
        g = self.cfg.grid
        F = torch.zeros((48, g.nxg, g.nyg, self.ldz), dtype=torch.float64, device=self.device)
        _lib.check(self.L.wamp3_compute_fluxes(self.ctx, self._u,
                                               _lib.ptrs48([F[m].data_ptr() for m in range(48)]),
                                               self.stream), "wamp3_compute_fluxes")
        return F[..., : g.nzg].cpu().numpy().copy()

    def run(self, nsteps: int):
        """n steps with dt from the state before each step; returns the dts."""
        torch = _torch()
        dts = torch.zeros(nsteps, dtype=torch.float64, device=self.device)
        self.compute_dt()
        for k in range(nsteps):
            dts[k].copy_(self.dt_buf[0])
            self.step_device()
        return dts.cpu().numpy()


class Batch3:
    """Slots of resident 3D cases sharing one scheme and precision (libwamp's
    wamp3_batch_*): each kernel of a stage is one launch over every active
    case, and each case's dt, time, end time and done flag stay on the device.
    Every case advances bit for bit as Solver3 does on it alone."""

    def __init__(self, scheme, precision: int, max_cases: int, device: str = "cuda"):
        torch = _torch()
        if not torch.cuda.is_available():
            raise _lib.WampError("batches need a ROCm GPU")
        self.L = _lib.load()
        self.device = torch.device(device)
        self.scheme = scheme
        self.precision = int(precision)
        self.max_cases = int(max_cases)
        sch = _lib.WampScheme(scheme.base, scheme.shear, scheme.vf_thinc, scheme.face_pp,
                              1 if scheme.repair else 0, scheme.cfl)
        h = C.c_void_p()
        _lib.check(self.L.wamp3_batch_create(C.byref(sch), self.precision, self.max_cases,
                                             C.byref(h)), "wamp3_batch_create")
        self.h = h
        self.cfgs: dict[int, Case3Config] = {}
        self.states: dict[int, object] = {}
        self.graph_steps = 0

    @property
    def stream(self) -> int:
        return _torch().cuda.current_stream(self.device).cuda_stream

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.wamp3_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_case(self, slot: int, cfg: Case3Config, U0, t0: float = 0.0, t_end: float = 0.0):
        torch = _torch()
        g = cfg.grid
        if cfg.precision != self.precision or cfg.scheme != self.scheme:
            raise ValueError("a batch holds one scheme and one precision")
        U = torch.zeros((7, g.nxg, g.nyg, g.nzg), dtype=torch.float64, device=self.device)
        src = U0 if isinstance(U0, torch.Tensor) else torch.from_numpy(
            np.ascontiguousarray(U0, dtype=np.float64))
        U.copy_(src[..., : g.nzg])
        grid = _lib.WampGrid3(g.nx, g.ny, g.nz, g.g, g.nzg, g.dx, g.dy, g.dz, (C.c_int32 * 6)(*g.bc))
        eos = _lib.WampEos(*cfg.eos.as_tuple())
        ptrs = _lib.PtrArr7(*[U[m].data_ptr() for m in range(7)])
        _lib.check(self.L.wamp3_batch_set_case(self.h, int(slot), C.byref(grid), C.byref(eos), ptrs,
                                               float(t0), float(t_end)), "wamp3_batch_set_case")
        self.cfgs[slot] = cfg
        self.states[slot] = U

    def clear_case(self, slot: int):
        _lib.check(self.L.wamp3_batch_clear_case(self.h, int(slot)), "wamp3_batch_clear_case")
        self.cfgs.pop(slot, None)
        self.states.pop(slot, None)

    def state(self, slot: int) -> np.ndarray:
        return self.states[slot].cpu().numpy().copy()

    def commit(self, graph_steps: int = 0):
        _lib.check(self.L.wamp3_batch_commit(self.h, self.stream), "wamp3_batch_commit")
        self.graph_steps = 0
        if graph_steps > 0:
            self.capture(graph_steps)

    def capture(self, steps: int):
        _torch().cuda.synchronize(self.device)
        _lib.check(self.L.wamp3_batch_graph(self.h, int(steps)), "wamp3_batch_graph")
        self.graph_steps = int(steps)

    def prepare(self, slots=None):
        if slots is None:
            _lib.check(self.L.wamp3_batch_prepare(self.h, None, -1, self.stream), "wamp3_batch_prepare")
        else:
            arr = (C.c_int32 * len(slots))(*slots)
            _lib.check(self.L.wamp3_batch_prepare(self.h, arr, len(slots), self.stream),
                       "wamp3_batch_prepare")

    def step(self) -> int:
        rc = self.L.wamp3_batch_step(self.h, self.stream)
        if rc <= 0:
            _lib.check(rc if rc != 0 else -1, "wamp3_batch_step")
        return rc

    def status(self):
        n = self.max_cases
        t, dt, done = (C.c_double * n)(), (C.c_double * n)(), (C.c_int32 * n)()
        _lib.check(self.L.wamp3_batch_status(self.h, t, dt, done, self.stream), "wamp3_batch_status")
        return np.array(t), np.array(dt), np.array(done, dtype=bool)


def main(argv=None) -> int:
    """python -m wamp.threed: run the 3D shock-sphere case and write npz outputs."""
    import argparse
    import os
    import sys
    import time

    from .config import SCHEMES
    from .io import infer_restart_step, output_name

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--n", type=int, default=64, help="grid 2n x n x n")
    ap.add_argument("--scheme", choices=sorted(SCHEMES), default="m10")
    ap.add_argument("--precision", type=int, choices=(0, 1), default=0)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--save-every", type=int, default=0, help="also write every K steps (0: last only)")
    ap.add_argument("--out", default="out3d")
    ap.add_argument("--restart", default=None, help="npz written by this program")
    args = ap.parse_args(argv)
    cfg, U0 = shock_sphere(args.n, scheme=SCHEMES[args.scheme])
    cfg = Case3Config(cfg.grid, cfg.eos, cfg.scheme, cfg.name, args.precision)
    t, step0 = 0.0, 0
    if args.restart:
        U0, t = load_restart3(args.restart, cfg)
        step0 = infer_restart_step(args.restart)
    os.makedirs(args.out, exist_ok=True)
    s = Solver3(cfg)
    s.set_state(U0)
    s.prepare()
    prefix = os.path.join(args.out, cfg.name)
    t0 = time.perf_counter()
    dts = []
    done = 0
    while done < args.steps:
        n = min(args.steps - done, args.save_every or args.steps)
        dts.extend(s.run(n))
        done += n
        t_now = t + float(np.sum(dts))
        path = save_npz3(output_name(prefix, step0 + done), cfg.grid, s.state(), t_now, cfg.eos)
        print(f"step {step0 + done}  t {t_now:.6g}  wrote {path}", file=sys.stderr, flush=True)
    el = time.perf_counter() - t0
    g = cfg.grid
    print(f"{args.steps} steps of {g.nx}x{g.ny}x{g.nz} in {el:.2f} s", file=sys.stderr, flush=True)
    s.close()
    return 0


if __name__ == "__main__":
    import sys

    sys.exit(main())
