"""Error measures between two solutions of the same case (mixed vs FP64).

Reported quantities: relative L1 error of density, pressure and volume
fraction over the interior, and the shift of the interface centroid
(sum a1*x / sum a1, same for y) in cells.
"""
from __future__ import annotations

import numpy as np

from .config import CaseConfig


def _pressure(cfg: CaseConfig, U):
    e = cfg.eos
    al = U[5]
    rho = U[0] + U[1]
    u, v = U[2] / rho, U[3] / rho
    a2 = 1.0 - al
    gb = al / (e.gamma1 - 1.0) + a2 / (e.gamma2 - 1.0)
    pb = al * e.gamma1 * e.pinf1 / (e.gamma1 - 1.0) + a2 * e.gamma2 * e.pinf2 / (e.gamma2 - 1.0)
    return (U[4] - pb - 0.5 * rho * (u * u + v * v)) / gb


def rel_l1(a, b) -> float:
    den = float(np.abs(b).sum())
    return float(np.abs(a - b).sum()) / (den if den > 0 else 1.0)


def centroid(cfg: CaseConfig, U):
    g = cfg.grid
    si, sj = g.interior()
    a = U[5][si, sj]
    x = g.x_centres()[si][:, None]
    y = g.y_centres()[sj][None, :]
    m = a.sum()
    return float((a * x).sum() / m), float((a * y).sum() / m)


def compare(cfg: CaseConfig, U, Uref) -> dict:
    """Errors of state U against the reference state Uref, both (6, NXG, NYG) FP64."""
    g = cfg.grid
    si, sj = g.interior()
    # interior cells only: the ghost layers may hold zeros (no density to divide by)
    Ui, Ri = U[:, si, sj], Uref[:, si, sj]
    rho, rho_r = Ui[0] + Ui[1], Ri[0] + Ri[1]
    p, p_r = _pressure(cfg, Ui), _pressure(cfg, Ri)
    cx, cy = centroid(cfg, U)
    rx, ry = centroid(cfg, Uref)
    return {
        "rho": rel_l1(rho, rho_r),
        "p": rel_l1(p, p_r),
        "alpha1": rel_l1(Ui[5], Ri[5]),
        "centroid_shift_cells": float(np.hypot((cx - rx) / g.dx, (cy - ry) / g.dy)),
        "finite": bool(np.isfinite(Ui).all()),
    }


def perturbed(cfg: CaseConfig, U0, eps: float, seed: int = 0):
    """U0 with the partial densities and the energy of every interior cell scaled
    by 1 + eps or 1 - eps (seeded signs): a round-off-sized perturbation."""
    g = cfg.grid
    si, sj = g.interior()
    U = np.array(U0, dtype=np.float64, copy=True)
    rng = np.random.default_rng(seed)
    for m in (0, 1, 4):
        sgn = rng.choice([-1.0, 1.0], size=U[m][si, sj].shape)
        U[m][si, sj] *= 1.0 + eps * sgn
    return U


def paired_run(cfg: CaseConfig, U0, steps: int, checkpoints=(), log=None, variant: str = "mixed",
               eps: float = 2.0 ** -52, t_end: float | None = None) -> list:
    """Advance one FP64 solver and one variant from the same state with the FP64
    run's dt sequence (so both sit at the same time); errors of the variant's
    state against the FP64 state at each checkpoint step.

    variant "mixed": the mixed-precision solver.  variant "perturbed": FP64 again,
    from the initial state perturbed by a relative eps per cell (`perturbed`),
    which measures how much the flow itself amplifies round-off-sized
    differences.  With t_end, stepping ends when the FP64 run reaches t_end
    (the last dt clipped), or after `steps` steps, whichever comes first."""
    import dataclasses

    import torch

    from .solver import Solver

    a = Solver(dataclasses.replace(cfg, precision=0))
    b = Solver(dataclasses.replace(cfg, precision=1 if variant == "mixed" else 0))
    a.set_state(U0)
    b.set_state(U0 if variant == "mixed" else perturbed(cfg, U0, eps))
    for s in (a, b):
        s.prepare()
    a.compute_dt()
    t = 0.0
    marks = sorted(set(int(c) for c in checkpoints) | {int(steps)})
    rows = []
    k = 0
    while k < int(steps):
        k += 1
        last = False
        if t_end is not None:
            dt = float(a.dt_buf[0].item())
            if t + dt >= t_end:
                a.dt_buf[0].fill_(t_end - t)
                last = True
        b.dt_buf[0].copy_(a.dt_buf[0])
        t += float(a.dt_buf[0].item())
        a.step_device()
        b.step_device()
        if k in marks or last:
            row = compare(cfg, b.state(), a.state())
            row.update(step=k, t=t)
            rows.append(row)
            if log:
                log(f"  step {k} (t {t:.4g}): rho {row['rho']:.3e}  p {row['p']:.3e}  "
                    f"alpha1 {row['alpha1']:.3e}  centroid {row['centroid_shift_cells']:.3e} cells")
        if last:
            break
    a.close()
    b.close()
    return rows


def main(argv=None) -> int:
    """Mixed-vs-FP64 (or perturbed-vs-FP64) error report for one named case (JSON on stdout)."""
    import argparse
    import json
    import sys

    from . import cli
    from .config import SCHEMES

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--case", choices=sorted(cli.CASES), default="multi_droplet")
    ap.add_argument("--nx", type=int, default=None)
    ap.add_argument("--ny", type=int, default=None)
    ap.add_argument("--scheme", choices=sorted(SCHEMES), default=None)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--checkpoints", type=str, default="1,10,100")
    ap.add_argument("--variant", choices=("mixed", "perturbed"), default="mixed",
                    help="compare mixed precision, or FP64 from a perturbed initial state")
    ap.add_argument("--eps", type=float, default=2.0 ** -52,
                    help="relative perturbation of the initial state (variant perturbed)")
    ap.add_argument("--t-end", type=float, default=None)
    args = ap.parse_args(argv)
    make, (nx0, ny0), scheme0, _ = cli.CASES[args.case]
    scheme = args.scheme or scheme0
    cfg, U0 = make(args.nx or nx0, args.ny or ny0, scheme=SCHEMES[scheme])
    cps = [int(c) for c in args.checkpoints.split(",") if c.strip()]
    rows = paired_run(cfg, U0, args.steps, cps, log=lambda m: print(m, file=sys.stderr, flush=True),
                      variant=args.variant, eps=args.eps, t_end=args.t_end)
    out = {"case": cfg.name, "scheme": scheme, "grid": [cfg.grid.nx, cfg.grid.ny],
           "variant": args.variant, "dt_sequence": "fp64 run's", "errors": rows}
    if args.variant == "perturbed":
        out["eps"] = args.eps
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    import sys

    sys.exit(main())
