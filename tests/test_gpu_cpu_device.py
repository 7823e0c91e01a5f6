"""The CPU device (wamp_cpu_*) against the FP64 GPU step on the same inputs:
the same per-cell source compiled for the host and for gfx950 gives the same
bits -- face fluxes, then tens of steps of state and dt, with both GPU
limiter/update kernels (LDS tile and register march), and a few steps of a
grid above 2^18 cells, where the march is the default."""
import numpy as np
import pytest

from wamp import cases
from wamp.config import SCHEMES

pytestmark = pytest.mark.gpu

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
        assert np.array_equal(x, y), (name, grp, float(np.abs(x - y).max()))
    gpu.close()
    cpu.close()


@pytest.mark.parametrize("lu_kind", ["0", "1"])
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
    a, b = gpu.state()[:, si, sj], cpu.state()[:, si, sj]
    assert np.array_equal(a, b), float(np.abs(a - b).max())
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
    assert np.array_equal(gpu.state()[:, si, sj], cpu.state()[:, si, sj])
    gpu.close()
    cpu.close()
