"""Script entry point with the reference's CLI (SC:13-21) and time loop (SC:1997-2082).

    python wave-multiphase_amd/wamp/cli.py [--cpu [--threads N] | --cuda]
        [--restart FILE.npz] [--restart-time T] [--restart-step N]
        [--case shock_cylinder|mach10_cylinder|multi_droplet|triple_point]
        [--nx NX --ny NY] [--scheme sc|m10|mp5] [--t-end T] [--save-times T1,T2,...]
        [--max-steps N] [--out DIR] [--prefix NAME] [--png]

The reference's flags keep their meaning: --cuda selects the GPU (the
default), --cpu the host (wamp.cpu.CpuSolver: the FP64 step's own code
compiled for the host, on --threads worker threads), and
--restart/--restart-time/--restart-step restart from an npz file.  --cpu
computes in FP64, as the reference's CPU run does.  Everything the reference
fixes in its source (grid, case, scheme, end time, save times) is a flag
here.  Outputs are `{prefix}_{step:06d}.npz` as SC:1714 / M10:1636.
"""
from __future__ import annotations

import argparse
import os
import sys
import time as _time

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from wamp import cases, io  # noqa: E402
from wamp.config import SCHEMES  # noqa: E402

CASES = {
    "shock_cylinder": (cases.shock_cylinder, (256, 128), "mp5", 0.12),
    "mach10_cylinder": (cases.mach10_cylinder, (4096, 2048), "m10", 0.02),
    "multi_droplet": (lambda nx, ny, scheme: cases.multi_droplet(nx, ny, scheme=scheme)[:2],
                      (4096, 2048), "mp5", 0.1),
    "triple_point": (cases.triple_point, (448, 192), "mp5", 1.0),
}


def parse(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--cpu", action="store_true")
    ap.add_argument("--cuda", action="store_true")
    ap.add_argument("--threads", type=int, default=0,
                    help="worker threads of the CPU device (--cpu); 0: OMP_NUM_THREADS when "
                         "set, else the CPUs this process may run on")
    ap.add_argument("--restart", type=str, default="")
    ap.add_argument("--restart-time", type=float, default=None)
    ap.add_argument("--restart-step", type=int, default=None)
    ap.add_argument("--case", choices=sorted(CASES), default="shock_cylinder")
    ap.add_argument("--nx", type=int, default=None)
    ap.add_argument("--ny", type=int, default=None)
    ap.add_argument("--scheme", choices=sorted(SCHEMES), default=None)
    ap.add_argument("--t-end", type=float, default=None)
    ap.add_argument("--save-times", type=str, default="")
    ap.add_argument("--max-steps", type=int, default=None)
    ap.add_argument("--out", type=str, default=".")
    ap.add_argument("--prefix", type=str, default=None)
    ap.add_argument("--log-every", type=int, default=200)
    ap.add_argument("--png", action="store_true",
                    help="also write {prefix}_{step:06d}.png (density gradient + alpha1) with the "
                         "final output, as SC:2080")
    args, _ = ap.parse_known_args(argv)
    return args


def main(argv=None) -> int:
    args = parse(argv)
    if args.cpu and args.cuda:
        print("--cpu and --cuda select different devices; give one", file=sys.stderr)
        return 2
    make, (nx0, ny0), scheme0, t_end0 = CASES[args.case]
    scheme_name = args.scheme or scheme0
    nx, ny = args.nx or nx0, args.ny or ny0
    cfg, U0 = make(nx, ny, scheme=SCHEMES[scheme_name])
    t_end = args.t_end if args.t_end is not None else t_end0
    saves = [float(x) for x in args.save_times.split(",") if x.strip()]
    prefix = args.prefix or ("mach10" if scheme_name == "m10" else "mp5")
    os.makedirs(args.out, exist_ok=True)
    g = cfg.grid
    if args.cpu:
        from wamp.cpu import CpuSolver

        s = CpuSolver(cfg, threads=args.threads)
        device = f"cpu ({s.threads} threads)"
    else:
        from wamp.solver import Solver

        s = Solver(cfg)
        device = "cuda"
    print(f"Device: {device}  NX={g.nx}  NY={g.ny}  G={g.g}  case={args.case}  scheme={scheme_name}")

    restart = args.restart != ""
    if restart:
        U, t_file = io.load_restart(args.restart, cfg)
        s.set_state(U)
        s.time = t_file if args.restart_time is None else float(args.restart_time)
        s.nstep = io.infer_restart_step(args.restart) if args.restart_step is None else int(args.restart_step)
        print(f"Restart: {args.restart}  t={s.time:.6e}  step={s.nstep}")
    else:
        s.set_state(U0)
    s.prepare()

    def save(sv):
        path = os.path.join(args.out, io.output_name(prefix, sv.nstep))
        io.save_npz(path, g, sv.state(), sv.prims(), sv.time)
        print(f"  saved {path}  t={sv.time:.2e}")

    if not restart:
        save(s)
    t0 = _time.time()
    s.advance_to(t_end, saves, on_save=save, log_every=args.log_every, max_steps=args.max_steps)
    save(s)
    if args.png:
        path = os.path.join(args.out, io.output_name(prefix, s.nstep, "png"))
        io.save_png(path, g, s.state())
        print(f"  saved {path}")
    print(f"Done. Total steps={s.nstep}, t={s.time:.4e}  wall={_time.time() - t0:.1f}s")
    s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
