"""Initial conditions for the solver's workloads (host-side, numpy, FP64).

The reference's own initial condition (ic_kernel2, SC:779-837 / M10:736-790)
is a dimensional sharp-interface setup; these are the nondimensional cases
the solver is measured on:

  shock_cylinder   Mach 2.4 shock on a water cylinder, 256x128 on [0,4]x[0,2]
  mach10_cylinder  Mach 10 shock on a water cylinder, 4096x2048 on [0,4]x[0,2]
  multi_droplet    Mach 2.4 shock on 40 seeded droplets, 25 % with air cavities
  batch_case       one of the uneven shock-droplet cases of a batch
  triple_point     the gas-gas triple point on [0,7]x[0,3]

States are returned as a (6, NXG, NYG) float64 array in the reference's
conserved layout [a1*rho1, a2*rho2, rho*u, rho*v, E, a1] (SC:777, 1842-1844),
ghost cells zero (the step's first boundary fill sets them).
"""
from __future__ import annotations

import math

import numpy as np

from .config import (BC_SLIP_WALL, BC_ZERO_GRADIENT, EOS, SCHEME_M10, SCHEME_MP5,
                     SCHEME_SC, CaseConfig, Grid)

ALPHA_CLIP = 1.0e-6
WATER_ND = EOS(gamma1=4.4, pinf1=6000.0, gamma2=1.4, pinf2=0.0)
TRIPLE_EOS = EOS(gamma1=1.5, pinf1=0.0, gamma2=1.4, pinf2=0.0)


def rankine_hugoniot(mach: float, gamma: float = 1.4, rho1: float = 1.0, p1: float = 1.0):
    """Post-shock (rho, p, u) behind a shock of Mach `mach` moving +x into gas at rest."""
    m2 = mach * mach
    rho2 = rho1 * (gamma + 1.0) * m2 / ((gamma - 1.0) * m2 + 2.0)
    p2 = p1 * (1.0 + 2.0 * gamma / (gamma + 1.0) * (m2 - 1.0))
    s = mach * math.sqrt(gamma * p1 / rho1)
    u2 = s * (1.0 - rho1 / rho2)
    return rho2, p2, u2


def conserved_from_prims(eos: EOS, a1, rho1, rho2, u, v, p):
    """[ar1, ar2, ru, rv, E, al] from primitive fields, E as prim_to_energy (SC:105-110)."""
    ar1 = a1 * rho1
    ar2 = (1.0 - a1) * rho2
    rho = ar1 + ar2
    a2 = 1.0 - a1
    gb = a1 / (eos.gamma1 - 1.0) + a2 / (eos.gamma2 - 1.0)
    pb = a1 * eos.gamma1 * eos.pinf1 / (eos.gamma1 - 1.0) + a2 * eos.gamma2 * eos.pinf2 / (eos.gamma2 - 1.0)
    E = gb * p + pb + 0.5 * rho * (u * u + v * v)
    return np.stack([ar1, ar2, rho * u, rho * v, E, a1]).astype(np.float64)


def _smeared(d, radius, width):
    return 0.5 * (1.0 - np.tanh((d - radius) / width))


def shock_droplets(grid: Grid, eos: EOS, mach: float, shock_x: float, droplets,
                   rho_water: float = 1000.0, rho_air: float = 1.0, p0: float = 1.0,
                   smear_cells: float = 1.5) -> np.ndarray:
    """Planar shock at `shock_x` moving +x into air at rest holding water droplets.

    `droplets` is a sequence of (xc, yc, r, r_cavity) with r_cavity = 0 for a
    solid droplet.  The interface is a tanh profile over `smear_cells` cells,
    alpha clipped to [1e-6, 1-1e-6].
    """
    xs = grid.x_centres()
    ys = grid.y_centres()
    x = xs[:, None]
    w = smear_cells * grid.dx
    water = np.zeros((grid.nxg, grid.nyg))
    for (xc, yc, r, rc) in droplets:
        # outside r + 12 w the profile is below 1e-10, i.e. lost in the alpha clip
        reach = r + 12.0 * w
        i0, i1 = np.searchsorted(xs, xc - reach), np.searchsorted(xs, xc + reach)
        j0, j1 = np.searchsorted(ys, yc - reach), np.searchsorted(ys, yc + reach)
        d = np.sqrt((xs[i0:i1, None] - xc) ** 2 + (ys[None, j0:j1] - yc) ** 2)
        a = _smeared(d, r, w)
        if rc > 0.0:
            a = a * (1.0 - _smeared(d, rc, w))
        np.maximum(water[i0:i1, j0:j1], a, out=water[i0:i1, j0:j1])
    a1 = np.clip(water, ALPHA_CLIP, 1.0 - ALPHA_CLIP)
    rho_s, p_s, u_s = rankine_hugoniot(mach, eos.gamma2, rho_air, p0)
    post = np.broadcast_to(x < shock_x, a1.shape)
    rho2 = np.where(post, rho_s, rho_air)
    p = np.where(post, p_s, p0)
    u = np.where(post, u_s, 0.0)
    v = np.zeros_like(u)
    U = conserved_from_prims(eos, a1, np.full_like(a1, rho_water), rho2, u, v, p)
    return _zero_ghosts(U, grid)


def _zero_ghosts(U, grid: Grid):
    out = np.zeros_like(U)
    si, sj = grid.interior()
    out[:, si, sj] = U[:, si, sj]
    return out


# ---------------------------------------------------------------------------
# Named workloads
# ---------------------------------------------------------------------------

def shock_cylinder(nx=256, ny=128, scheme=SCHEME_MP5):
    """Mach 2.4 shock at x=0.8 on a water cylinder r=0.5 at (1.5, 1.0), [0,4]x[0,2]."""
    grid = Grid(nx, ny, 0.0, 4.0, 0.0, 2.0)
    cfg = CaseConfig(grid, WATER_ND, scheme, name=f"shock_cylinder_{nx}x{ny}")
    U = shock_droplets(grid, WATER_ND, 2.4, 0.8, [(1.5, 1.0, 0.5, 0.0)])
    return cfg, U


def mach10_cylinder(nx=4096, ny=2048, scheme=SCHEME_M10):
    """Mach 10 shock at x=0.6 on a water cylinder r=0.25 at (1.0, 1.0), [0,4]x[0,2]."""
    grid = Grid(nx, ny, 0.0, 4.0, 0.0, 2.0)
    cfg = CaseConfig(grid, WATER_ND, scheme, name=f"mach10_cylinder_{nx}x{ny}")
    U = shock_droplets(grid, WATER_ND, 10.0, 0.6, [(1.0, 1.0, 0.25, 0.0)])
    return cfg, U


def seeded_droplets(rng, n, rmin, rmax, xlo, xhi, ylo, yhi, gap, n_cavity,
                    cavity_frac=0.4, max_tries=200000):
    """Rejection-sampled non-overlapping droplets; `n_cavity` of them carry a cavity."""
    drops = []
    tries = 0
    while len(drops) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError(f"could not place {n} droplets (placed {len(drops)})")
        r = rng.uniform(rmin, rmax)
        xc = rng.uniform(xlo + r, xhi - r)
        yc = rng.uniform(ylo + r, yhi - r)
        if all((xc - a) ** 2 + (yc - b) ** 2 >= (r + rr + gap) ** 2 for (a, b, rr) in drops):
            drops.append((xc, yc, r))
    cav = set(rng.choice(n, size=n_cavity, replace=False).tolist()) if n_cavity else set()
    return [(xc, yc, r, cavity_frac * r if k in cav else 0.0) for k, (xc, yc, r) in enumerate(drops)]


def multi_droplet(nx=4096, ny=2048, n_drops=40, seed=3, scheme=SCHEME_MP5):
    """Mach 2.4 shock at x=0.8 on `n_drops` seeded droplets, r in [0.05, 0.15], [0,4]x[0,2]."""
    grid = Grid(nx, ny, 0.0, 4.0, 0.0, 2.0)
    rng = np.random.default_rng(seed)
    drops = seeded_droplets(rng, n_drops, 0.05, 0.15, 0.9, 3.9, 0.05, 1.95,
                            4.0 * grid.dx, n_cavity=n_drops // 4)
    cfg = CaseConfig(grid, WATER_ND, scheme, name=f"multi_droplet_{nx}x{ny}")
    return cfg, shock_droplets(grid, WATER_ND, 2.4, 0.8, drops), drops


BATCH_GRIDS = [(256, 256), (384, 256), (512, 256), (768, 384), (1024, 512)]


def batch_case(rng, scheme=SCHEME_MP5, grids=BATCH_GRIDS):
    """One uneven shock-droplet case: grid from `grids`, Mach U[1.5,6], r U[0.2,0.5]."""
    nx, ny = grids[int(rng.integers(len(grids)))]
    ly = 2.0
    lx = ly * nx / ny
    grid = Grid(nx, ny, 0.0, lx, 0.0, ly)
    mach = float(rng.uniform(1.5, 6.0))
    r = float(rng.uniform(0.2, 0.5))
    # the shock sits at x = xc - r - 0.2 and must leave room for 0.1 of post-shock gas
    xc = float(rng.uniform(r + 0.4, lx - r - 0.1))
    yc = float(rng.uniform(r + 0.1, ly - r - 0.1))
    cfg = CaseConfig(grid, WATER_ND, scheme, name=f"batch_{nx}x{ny}_M{mach:.2f}")
    U = shock_droplets(grid, WATER_ND, mach, xc - r - 0.2, [(xc, yc, r, 0.0)])
    return cfg, U


def triple_point_grid(nx, ny) -> Grid:
    """[0,7]x[0,3], all walls slip."""
    return Grid(nx, ny, 0.0, 7.0, 0.0, 3.0,
                bc=(BC_SLIP_WALL, BC_SLIP_WALL, BC_SLIP_WALL, BC_SLIP_WALL))


def triple_point(nx=448, ny=192, scheme=SCHEME_MP5):
    """Gas-gas triple point on [0,7]x[0,3], all walls slip."""
    grid = triple_point_grid(nx, ny)
    x = grid.x_centres()[:, None] * np.ones((1, grid.nyg))
    y = grid.y_centres()[None, :] * np.ones((grid.nxg, 1))
    left = x < 1.0
    top = (~left) & (y >= 1.5)
    a1 = np.where(left | top, 1.0 - ALPHA_CLIP, ALPHA_CLIP)
    rho = np.where(left, 1.0, np.where(top, 0.125, 1.0))
    p = np.where(left, 1.0, 0.1)
    z = np.zeros_like(x)
    U = conserved_from_prims(TRIPLE_EOS, a1, rho, rho, z, z, p)
    cfg = CaseConfig(grid, TRIPLE_EOS, scheme, name=f"triple_point_{nx}x{ny}")
    return cfg, _zero_ghosts(U, grid)


# ---------------------------------------------------------------------------
# The same cases built in device memory (wamp_init_droplets / _triple_point)
# ---------------------------------------------------------------------------

def _abi_grid_eos(grid: Grid, eos: EOS, ld: int):
    import ctypes as C

    from . import _lib

    return (_lib.WampGrid(grid.nx, grid.ny, grid.g, ld, grid.dx, grid.dy, (C.c_int32 * 4)(*grid.bc)),
            _lib.WampEos(*eos.as_tuple()))


def shock_droplets_device(grid: Grid, eos: EOS, mach: float, shock_x: float, droplets,
                          rho_water: float = 1000.0, rho_air: float = 1.0, p0: float = 1.0,
                          smear_cells: float = 1.5, device: str = "cuda"):
    """Device twin of shock_droplets: a (6, NXG, NYG) FP64 tensor built by one kernel.
    Equal to the host state up to the last-ulp differences of tanh between the
    device and host math libraries."""
    import ctypes as C

    import torch

    from . import _lib

    L = _lib.load()
    ld = grid.nyg
    U = torch.empty((6, grid.nxg, ld), dtype=torch.float64, device=device)
    drops = np.asarray(droplets, dtype=np.float64).reshape(-1, 4)
    d_dev = torch.from_numpy(drops).to(device) if len(drops) else None
    rho_s, p_s, u_s = rankine_hugoniot(mach, eos.gamma2, rho_air, p0)
    ic = _lib.WampDropletIC(grid.xmin, grid.ymin, shock_x, rho_s, p_s, u_s, rho_water, rho_air, p0,
                            smear_cells * grid.dx, d_dev.data_ptr() if d_dev is not None else None,
                            len(drops))
    g_s, e_s = _abi_grid_eos(grid, eos, ld)
    stream = torch.cuda.current_stream(U.device).cuda_stream
    _lib.check(L.wamp_init_droplets(C.byref(g_s), C.byref(e_s), C.byref(ic),
                                    _lib.ptrs6([U[m].data_ptr() for m in range(6)]), stream),
               "wamp_init_droplets")
    return U  # d_dev is freed in stream order after the kernel


def triple_point_device(nx=448, ny=192, scheme=SCHEME_MP5, device: str = "cuda"):
    """Device twin of triple_point (bitwise the same state)."""
    import ctypes as C

    import torch

    from . import _lib

    L = _lib.load()
    grid = triple_point_grid(nx, ny)
    U = torch.empty((6, grid.nxg, grid.nyg), dtype=torch.float64, device=device)
    ic = _lib.WampTripleIC(grid.xmin, grid.ymin, 1.0, 1.5, 1.0, 1.0, 0.125, 0.1, 1.0, 0.1)
    g_s, e_s = _abi_grid_eos(grid, TRIPLE_EOS, grid.nyg)
    stream = torch.cuda.current_stream(U.device).cuda_stream
    _lib.check(L.wamp_init_triple_point(C.byref(g_s), C.byref(e_s), C.byref(ic),
                                        _lib.ptrs6([U[m].data_ptr() for m in range(6)]), stream),
               "wamp_init_triple_point")
    cfg = CaseConfig(grid, TRIPLE_EOS, scheme, name=f"triple_point_{nx}x{ny}")
    return cfg, U


__all__ = [
    "rankine_hugoniot", "conserved_from_prims", "shock_droplets", "shock_cylinder",
    "mach10_cylinder", "multi_droplet", "seeded_droplets", "batch_case", "triple_point",
    "WATER_ND", "TRIPLE_EOS", "BATCH_GRIDS", "SCHEME_SC", "SCHEME_M10", "SCHEME_MP5",
    "BC_ZERO_GRADIENT", "shock_droplets_device", "triple_point_device",
]
