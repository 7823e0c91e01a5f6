"""The CPU device (wamp_cpu_*) against the FP64 GPU step on the same inputs:
the same per-cell source compiled for the host and for gfx950 gives the same
bits -- face fluxes, then tens of steps of state and dt, with both GPU
limiter/update kernels (LDS tile and register march), and a few steps of a
grid above 2^18 cells, where the march is the default.

The one exception is subnormal values (below 2^-1022, about 2.2e-308): the
kernels' shortened division sequence is the IEEE quotient only for quotients
of at least 2^-1000, and the host divides by IEEE rules, so a momentum that
decays into the subnormal range (a fluid at rest next to a moving one) may
differ in its last subnormal bits.  Every normal value must be bit-identical."""
import numpy as np
import pytest

from wamp import cases
from wamp.config import SCHEMES

pytestmark = pytest.mark.gpu

TINY = np.finfo(np.float64).tiny  # smallest normal double


def assert_same_but_subnormal(a, b, what=""):
    """a and b bit-identical wherever either is a normal number."""
    diff = a != b
    normal = (np.abs(a) >= TINY) | (np.abs(b) >= TINY)
    bad = diff & normal
    assert not bad.any(), (what, int(bad.sum()), float(np.abs(a - b)[bad].max()))
    if diff.any():
        print(f"{what}: {int(diff.sum())} subnormal values differ, max {float(np.abs(a - b).max()):.3e}")


CASES = {
    "sc": lambda: cases.shock_cylinder(96, 48),
    "m10": lambda: cases.mach10_cylinder(128, 64),
    "mp5": lambda: cases.shock_cylinder(128, 64, scheme=SCHEMES["mp5"]),
    "triple": lambda: cases.triple_point(112, 48),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_cpu_device_fluxes_equal_gpu(gpu_ok, name):
    from wamp.cpu import CpuSolver
    from wamp.solver import Solver

    cfg, U0 = CASES[name]()
    gpu = Solver(cfg)
    gpu.set_state(U0)
    gpu.prepare()
    cpu = CpuSolver(cfg)
    cpu.set_state(U0)
    cpu.prepare()
    a, b = gpu.fluxes(), cpu.fluxes()
    g = cfg.grid
    xs = (slice(g.g - 1, g.g + g.nx), slice(g.g, g.g + g.ny))
    ys = (slice(g.g, g.g + g.nx), slice(g.g - 1, g.g + g.ny))
    for grp, (r, c) in enumerate((xs, ys, xs, ys)):
        x, y = a[7 * grp: 7 * grp + 7, r, c], b[7 * grp: 7 * grp + 7, r, c]
        assert np.isfinite(y).all()
        assert_same_but_subnormal(x, y, f"{name} fluxes {grp}")
    gpu.close()
    cpu.close()


@pytest.mark.parametrize("lu_kind", ["0", "1", "2"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_cpu_device_steps_equal_gpu(gpu_ok, name, lu_kind, monkeypatch):
    from wamp.cpu import CpuSolver
    from wamp.solver import Solver

    monkeypatch.setenv("WAMP_LU_KIND", lu_kind)
    cfg, U0 = CASES[name]()
    gpu = Solver(cfg)
    gpu.set_state(U0)
    gpu.prepare()
    cpu = CpuSolver(cfg)
    cpu.set_state(U0)
    cpu.prepare()
    dg, dc = gpu.run(40), cpu.run(40)
    assert np.array_equal(dg, dc), (dg, dc)
    si, sj = cfg.grid.interior()
    assert_same_but_subnormal(gpu.state()[:, si, sj], cpu.state()[:, si, sj], f"{name} state")
    gpu.close()
    cpu.close()


def test_cpu_device_equals_gpu_above_the_march_threshold(gpu_ok):
    from wamp.cpu import CpuSolver
    from wamp.solver import Solver

    cfg, U0 = cases.mach10_cylinder(768, 384)  # 294,912 cells
    gpu = Solver(cfg)
    gpu.set_state(U0)
    gpu.prepare()
    cpu = CpuSolver(cfg)
    cpu.set_state(U0)
    cpu.prepare()
    dg, dc = gpu.run(4), cpu.run(4)
    assert np.array_equal(dg, dc)
    si, sj = cfg.grid.interior()
    assert_same_but_subnormal(gpu.state()[:, si, sj], cpu.state()[:, si, sj], "768x384 state")
    gpu.close()
    cpu.close()
