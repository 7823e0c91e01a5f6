"""Write the oracle-generated regression fixtures in tests/golden/.

These are NOT reference-held vectors (the reference ships none): they are
outputs of the CPU oracle (oracle/wamp_oracle.c) on small shock-cylinder
cases, committed so that later changes to the oracle or to the HIP path are
checked against a fixed anchor.  Regenerate only on a deliberate oracle
change:

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "wave-multiphase_amd")]

from oracle.oracle import Oracle  # noqa: E402
from wamp import cases  # noqa: E402
from wamp.config import SCHEMES  # noqa: E402

NX, NY, STEPS = 32, 16, 6


def make(scheme: str):
    cfg, U0 = cases.shock_cylinder(NX, NY, scheme=SCHEMES[scheme])
    orc = Oracle(cfg, threads=1)
    orc.set_state(U0)
    orc.prepare()
    dts = orc.run(STEPS)
    si, sj = cfg.grid.interior()
    S = orc.state()[:, si, sj]
    path = os.path.join(HERE, f"shock_cylinder_{NX}x{NY}_{scheme}_{STEPS}steps.npz")
    np.savez_compressed(path, state=S, dts=dts, nx=NX, ny=NY, steps=STEPS, scheme=scheme)
    return path


if __name__ == "__main__":
    for s in ("sc", "m10", "mp5"):
        print(make(s))
