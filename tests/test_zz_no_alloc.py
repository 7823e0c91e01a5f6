"""wamp_step_noalloc, the step that allocates nothing (its scratch is a static
device array): bit for bit the result of wamp_step on the same state, call
after call on one device with grids of different sizes (each call reuses the
array the previous one left behind), in both precisions; grids whose scratch
does not fit are refused with a message before anything runs."""
import ctypes as C
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _abi(cfg, ld):
    from wamp import _lib

    g, s = cfg.grid, cfg.scheme
    return (_lib.WampGrid(g.nx, g.ny, g.g, ld, g.dx, g.dy, (C.c_int32 * 4)(*g.bc)),
            _lib.WampEos(*cfg.eos.as_tuple()),
            _lib.WampScheme(s.base, s.shear, s.vf_thinc, s.face_pp, 1 if s.repair else 0, s.cfl))


def _one(cfg, U0):
    """(wamp_step result, wamp_step_noalloc result) from the same prepared state."""
    import torch

    from wamp import _lib
    from wamp.solver import Solver

    ref = Solver(cfg)
    ref.set_state(U0)
    ref.prepare()
    for _ in range(3):          # a shocked state
        ref.compute_dt()
        ref.step_device()
    ref.compute_dt()
    U = ref.U.clone()
    dt = ref.dt_buf.clone()
    ref.step_device()
    want_u, want_dt = ref.U.cpu().numpy(), float(ref.dt_buf[0].item())
    ref.close()
    grid, eos, sch = _abi(cfg, U.shape[2])
    st = torch.cuda.current_stream().cuda_stream
    _lib.check(ref.L.wamp_step_noalloc(C.byref(grid), C.byref(eos), C.byref(sch), int(cfg.precision),
                                       _lib.ptrs6([U[m].data_ptr() for m in range(6)]),
                                       dt.data_ptr(), dt.data_ptr() + 8, st), "wamp_step_noalloc")
    torch.cuda.synchronize()
    return (want_u, want_dt), (U.cpu().numpy(), float(dt[1].item()))


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_step_equals_wamp_step_call_after_call(gpu_ok, precision):
    from wamp import cases
    from wamp.config import SCHEME_M10, SCHEME_MP5

    # larger, then smaller, then the larger again on another scheme: each
    # call finds the previous call's contents in the static scratch
    seq = [cases.shock_cylinder(256, 128, scheme=SCHEME_M10),
           cases.shock_cylinder(96, 48, scheme=SCHEME_M10),
           cases.shock_cylinder(200, 120, scheme=SCHEME_MP5)]
    for cfg, U0 in seq:
        cfg = dataclasses.replace(cfg, precision=precision)
        (want_u, want_dt), (got_u, got_dt) = _one(cfg, U0)
        si, sj = cfg.grid.interior()
        assert np.array_equal(got_u[:, si, sj], want_u[:, si, sj]), (cfg.name, precision)
        assert got_dt == want_dt, (cfg.name, precision)


def test_noalloc_refuses_grids_beyond_its_scratch(gpu_ok):
    import torch

    from wamp import _lib, cases
    from wamp.config import SCHEME_M10

    L = _lib.load()
    cfg, _ = cases.mach10_cylinder(1024, 512, scheme=SCHEME_M10)
    grid, eos, sch = _abi(cfg, cfg.grid.nyg)
    dummy = torch.zeros(8, dtype=torch.float64, device="cuda")
    p = dummy.data_ptr()
    rc = L.wamp_step_noalloc(C.byref(grid), C.byref(eos), C.byref(sch), 0, _lib.ptrs6([p] * 6),
                             p, p, torch.cuda.current_stream().cuda_stream)
    assert rc < 0
    msg = L.wamp_last_error().decode()
    assert "static scratch holds" in msg and "1024 x 512" in msg
    assert torch.all(dummy == 0)    # nothing ran
