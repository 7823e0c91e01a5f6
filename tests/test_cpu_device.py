"""The CPU device (the reference's --cpu, SC:15-21): libwamp.so's wamp_cpu_*
entry points run the FP64 step's own per-cell code on the host.  Its results
are checked bit for bit against the oracle built with the same pow and tanh
(liborc_wm.so: the C restatement of the reference's kernels, whose other
arithmetic is IEEE operations in the reference's order): face fluxes, the
state and every dt over tens of steps, for the SC, M10 and MP5 schemes and the
gas-gas triple point; the boundary fill and cons_to_prim; invariance to the
thread count and the row stride; argument errors.  No GPU is used."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

from wamp import _lib, cases
from wamp.config import SCHEMES
from wamp.cpu import CpuSolver
from oracle.oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "wave-multiphase_amd", "wamp", "cli.py")

CASES = {
    "sc": lambda: cases.shock_cylinder(48, 24),
    "m10": lambda: cases.mach10_cylinder(64, 32),
    "mp5": lambda: cases.shock_cylinder(64, 32, scheme=SCHEMES["mp5"]),
    "triple": lambda: cases.triple_point(56, 24),
}


def _pair(name, threads=3):
    cfg, U0 = CASES[name]()
    orc = Oracle(cfg, shared_math=True)
    orc.set_state(U0)
    orc.prepare()
    cpu = CpuSolver(cfg, threads=threads)
    cpu.set_state(U0)
    cpu.prepare()
    return cfg, orc, cpu


def _face_regions(g):
    xs = (slice(g.g - 1, g.g + g.nx), slice(g.g, g.g + g.ny))
    ys = (slice(g.g, g.g + g.nx), slice(g.g - 1, g.g + g.ny))
    return [xs, ys, xs, ys]  # groups of 7: high x, high y, low x, low y


@pytest.mark.parametrize("name", sorted(CASES))
def test_fluxes_bitwise_equal_to_oracle(name):
    cfg, orc, cpu = _pair(name)
    orc.apply_bc()
    orc.update_prims()
    orc.apply_bc()
    orc.compute_fluxes()
    ref, got = orc.fluxes(), cpu.fluxes()
    for grp, (a, b) in enumerate(_face_regions(cfg.grid)):
        r, x = ref[7 * grp: 7 * grp + 7, a, b], got[7 * grp: 7 * grp + 7, a, b]
        assert np.isfinite(x).all()
        assert np.array_equal(r, x), (name, grp, float(np.abs(r - x).max()))


@pytest.mark.parametrize("name", sorted(CASES))
def test_steps_bitwise_equal_to_oracle(name):
    cfg, orc, cpu = _pair(name)
    assert cpu.get_dt() == orc.get_dt()
    d_ref, d_got = orc.run(30), cpu.run(30)
    assert np.array_equal(d_ref, d_got), (d_ref, d_got)
    si, sj = cfg.grid.interior()
    a, b = orc.state()[:, si, sj], cpu.state()[:, si, sj]
    assert np.array_equal(a, b), float(np.abs(a - b).max())
    # a step that moves the state (not a fixed point)
    cfg0, U0 = CASES[name]()
    assert not np.array_equal(U0[:, si, sj], b)


def test_boundary_fill_and_prims_bitwise_equal_to_oracle():
    cfg, orc, cpu = _pair("sc")
    orc.run(3)
    cpu.run(3)
    orc.apply_bc()
    cpu.apply_bc()
    assert np.array_equal(orc.state(), cpu.state())  # ghost layers included
    orc.update_prims()
    orc.apply_bc()
    P = cpu.prims()
    ref = orc.prims()
    assert np.array_equal(ref, P), float(np.abs(ref - P).max())


def test_thread_count_and_row_stride_do_not_change_the_bits():
    cfg, U0 = CASES["m10"]()
    runs = []
    for threads, ld in ((1, None), (4, None), (7, cfg.grid.nyg + 5)):
        s = CpuSolver(cfg, threads=threads, ld=ld)
        assert s.threads == threads
        s.set_state(U0)
        s.prepare()
        d = s.run(8)
        runs.append((d, s.state()))
        s.close()
    si, sj = cfg.grid.interior()
    for d, st in runs[1:]:
        assert np.array_equal(d, runs[0][0])
        assert np.array_equal(st[:, si, sj], runs[0][1][:, si, sj])


def test_time_loop_matches_the_oracle_to_the_save_times():
    cfg, orc, cpu = _pair("sc")
    saved = []
    t = cpu.advance_to(0.01, save_times=[0.004], on_save=lambda s: saved.append(s.time), log=None)
    assert t == pytest.approx(0.01) and len(saved) == 1
    # the same dt sequence, clipped the same way (SC:2019-2026)
    time, dt, flag = 0.0, orc.get_dt(), False
    while time < 0.01 - 1e-14:
        if not flag and time + dt > 0.004:
            dt = 0.004 - time
        if time + dt > 0.01:
            dt = 0.01 - time
        orc.step(dt)
        time += dt
        dt = orc.get_dt()
        if not flag and abs(time - 0.004) < 0.5 * dt:
            flag = True
    si, sj = cfg.grid.interior()
    assert time == t
    assert np.array_equal(orc.state()[:, si, sj], cpu.state()[:, si, sj])


def test_argument_errors():
    cfg, U0 = CASES["sc"]()
    with pytest.raises(_lib.WampError, match="FP64 only"):
        CpuSolver(cfg.with_precision(1))
    L = _lib.load()
    g = cfg.grid
    grid = _lib.WampGrid(g.nx, g.ny, g.g, g.nyg, g.dx, g.dy, (C.c_int32 * 4)(*g.bc))
    eos = _lib.WampEos(*cfg.eos.as_tuple())
    s = cfg.scheme
    sch = _lib.WampScheme(s.base, s.shear, s.vf_thinc, s.face_pp, 1, s.cfl)
    ctx = C.c_void_p()
    assert L.wamp_cpu_ctx_create(C.byref(grid), C.byref(eos), C.byref(sch), -1, C.byref(ctx)) < 0
    assert "threads" in L.wamp_last_error().decode() and not ctx.value
    bad = _lib.WampGrid(g.nx, g.ny, 3, g.nyg, g.dx, g.dy, (C.c_int32 * 4)(*g.bc))
    assert L.wamp_cpu_ctx_create(C.byref(bad), C.byref(eos), C.byref(sch), 1, C.byref(ctx)) < 0
    assert "g must be >= 4" in L.wamp_last_error().decode()
    cpu = CpuSolver(cfg, threads=2)
    cpu.set_state(U0)
    with pytest.raises(_lib.WampError, match="dt must be > 0"):
        cpu.step_dt(0.0)
    with pytest.raises(_lib.WampError, match="is NULL"):
        _lib.check(L.wamp_cpu_step(cpu.ctx, _lib.ptrs6([0] * 6), 1e-3, None), "wamp_cpu_step")
    F = np.zeros((28, g.nxg, g.nyg))
    ptrs = [F[m].ctypes.data for m in range(28)]
    ptrs[5] += 8
    with pytest.raises(_lib.WampError, match="equally spaced"):
        _lib.check(L.wamp_cpu_compute_fluxes(cpu.ctx, cpu._u, _lib.ptrs28(ptrs)), "fluxes")
    assert L.wamp_cpu_step(None, cpu._u, 1e-3, None) < 0


def test_cli_cpu_runs_and_restarts(tmp_path):
    """--cpu runs the case on the host, writes the outputs and restarts from them."""
    base = [sys.executable, CLI, "--cpu", "--threads", "2", "--case", "shock_cylinder", "--nx", "48",
            "--ny", "24", "--scheme", "sc", "--log-every", "0"]
    r = subprocess.run(base + ["--t-end", "0.01", "--save-times", "0.005", "--out", str(tmp_path / "a")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "Device: cpu (2 threads)" in r.stdout
    files = sorted(os.listdir(tmp_path / "a"))
    assert files[0] == "mp5_000000.npz" and len(files) == 3, files
    end = np.load(tmp_path / "a" / files[-1])
    assert float(end["time"]) == pytest.approx(0.01)
    for k in ("rho", "u", "v", "p", "alpha1"):
        assert np.isfinite(end[k]).all(), k
    mid = files[1]
    r = subprocess.run(base + ["--t-end", "0.01", "--restart", str(tmp_path / "a" / mid),
                               "--out", str(tmp_path / "b")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "Restart:" in r.stdout
    r = subprocess.run([sys.executable, CLI, "--cpu", "--cuda"], capture_output=True, text=True)
    assert r.returncode == 2 and "give one" in r.stderr
