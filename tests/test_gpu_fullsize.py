"""The headline case at its full size, 4096x2048 (Mach-10 water cylinder, M10):
one SSP-RK3 step of the HIP path against the CPU oracle, then size-independent
properties over many steps -- mirror symmetry about the cylinder's axis (the
case and the slip walls are symmetric in y), admissibility and finiteness."""
import os

import numpy as np
import pytest

from conftest import rel_l1

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    from wamp import cases
    from wamp.config import SCHEME_M10

    return cases.mach10_cylinder(4096, 2048, scheme=SCHEME_M10)


def test_full_size_step_matches_oracle(gpu_ok, case):
    from oracle.oracle import Oracle
    from wamp.solver import Solver

    cfg, U0 = case
    threads = int(os.environ.get("OMP_NUM_THREADS", "16"))
    orc = Oracle(cfg, threads=threads)
    orc.set_state(U0)
    orc.prepare()
    U = orc.state()
    dt = orc.get_dt()
    orc.step(dt)
    ref = orc.state()
    del orc
    s = Solver(cfg)
    s.set_state(U)
    s.set_dt(dt)
    s.step_device()
    got = s.state()
    s.close()
    si, sj = cfg.grid.interior()
    errs = [rel_l1(got[m][si, sj], ref[m][si, sj]) for m in range(6)]
    print("4096x2048 one step:", ["%.2e" % e for e in errs])
    assert max(errs) <= 1e-12


def test_full_size_symmetry_and_admissibility(gpu_ok, case):
    from wamp.solver import Solver

    cfg, U0 = case
    s = Solver(cfg)
    s.set_state(U0)
    s.prepare()
    dts = s.run(100)
    S = s.state()
    s.close()
    assert np.all(dts > 0) and np.all(np.isfinite(dts))
    si, sj = cfg.grid.interior()
    worst = 0.0
    for m in range(6):
        a = S[m][si, sj]
        assert np.isfinite(a).all(), m
        b = a[:, ::-1] * (-1.0 if m == 3 else 1.0)  # v changes sign under the mirror
        worst = max(worst, rel_l1(a, b) if m != 3 else float(np.abs(a - b).sum() / np.abs(S[2][si, sj]).sum()))
    print(f"4096x2048 after 100 steps: mirror asymmetry {worst:.2e}")
    assert worst <= 1e-10
    ar1, ar2, al = S[0][si, sj], S[1][si, sj], S[5][si, sj]
    assert ar1.min() > 0.0 and ar2.min() > 0.0
    assert al.min() >= 1.0e-8 and al.max() <= 1.0 - 1.0e-8


@pytest.mark.skipif(not os.environ.get("WAMP_FULLSIZE_STEPS"),
                    reason="set WAMP_FULLSIZE_STEPS=N to compare N full-size steps with the oracle "
                           "(the oracle takes ~2.3 s per step on 16 cores)")
def test_full_size_many_steps_match_oracle(gpu_ok, case):
    from oracle.oracle import Oracle
    from wamp.solver import Solver

    cfg, U0 = case
    n = int(os.environ["WAMP_FULLSIZE_STEPS"])
    orc = Oracle(cfg, threads=int(os.environ.get("OMP_NUM_THREADS", "16")))
    orc.set_state(U0)
    orc.prepare()
    U = orc.state()
    chunks = []
    for k0 in range(0, n, 25):  # the same run as orc.run(n), with progress lines
        chunks.append(orc.run(min(25, n - k0)))
        print(f"oracle: {k0 + len(chunks[-1])} of {n} steps", flush=True)
    dts_ref = np.concatenate(chunks)
    ref = orc.state()
    del orc
    s = Solver(cfg)
    s.set_state(U)
    dts = s.run(n)
    got = s.state()
    s.close()
    si, sj = cfg.grid.interior()
    errs = [rel_l1(got[m][si, sj], ref[m][si, sj]) for m in range(6)]
    print(f"4096x2048, {n} steps:", ["%.2e" % e for e in errs],
          "dt rel", float(np.abs(dts - dts_ref).max() / dts_ref.max()))
    assert max(errs) <= 1e-9


@pytest.mark.skipif(os.environ.get("WAMP_MAXSIZE") != "1",
                    reason="opt-in (WAMP_MAXSIZE=1): 16384x8192, ~80 GB of host memory, ~1 minute")
def test_largest_grid_step_matches_oracle(gpu_ok):
    """16384 x 8192 (134 M cells, 16x the headline): one step against the oracle
    -- 64-bit offsets, launch grids and the march's strip / segment counts at
    the size where 32-bit cell indices would no longer be safe for ld * nx."""
    from oracle.oracle import Oracle
    from wamp import cases
    from wamp.config import SCHEME_M10
    from wamp.solver import Solver

    cfg, U0 = cases.mach10_cylinder(16384, 8192, scheme=SCHEME_M10)
    orc = Oracle(cfg, threads=int(os.environ.get("OMP_NUM_THREADS", "16")))
    orc.set_state(U0)
    del U0
    orc.prepare()
    U = orc.state()
    dt = orc.get_dt()
    orc.step(dt)
    ref = orc.state()
    dt_ref = orc.get_dt()
    orc.close()
    s = Solver(cfg)
    s.set_state(U)
    s.set_dt(dt)
    s.step_device()
    got = s.state()
    dt_next = float(s.dt_buf[0].item())
    s.close()
    si, sj = cfg.grid.interior()
    errs = [rel_l1(got[m][si, sj], ref[m][si, sj]) for m in range(6)]
    print("16384x8192 one step:", ["%.2e" % e for e in errs])
    assert max(errs) <= 1e-12
    assert abs(dt_next - dt_ref) <= 1e-14 * dt_ref
