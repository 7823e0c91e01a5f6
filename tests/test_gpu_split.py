"""The split entry points -- wamp_compute_fluxes into caller-owned face planes,
then wamp_limit_update per RK stage, then wamp_get_dt -- reproduce wamp_step bit
for bit (the stage structure of SC:2028-2059 kept visible to a caller); the 3D
pair wamp3_compute_fluxes / wamp3_limit_update reproduces wamp3_step the same way."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lu_kind", ["0", "1", "2"])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("scheme", ["sc", "m10", "mp5"])
def test_split_calls_equal_one_step(gpu_ok, scheme, precision, lu_kind, monkeypatch):
    """In mixed precision wamp_step keeps the face fluxes in the mixed storage
    (FP32 planes, FP64 energy planes) while the split calls pass 28 FP64 planes:
    both limiter kernels (lu_kind 0: LDS tile, 1: march) must read the same
    numbers from either."""
    import dataclasses

    import torch

    from wamp import _lib, cases
    from wamp.config import SCHEMES
    from wamp.solver import Solver

    monkeypatch.setenv("WAMP_LU_KIND", lu_kind)
    cfg, U0 = cases.shock_cylinder(160, 80, scheme=SCHEMES[scheme])
    cfg = dataclasses.replace(cfg, precision=precision)
    ref = Solver(cfg)
    ref.set_state(U0)
    ref.prepare()
    for _ in range(15):  # a shocked, interacting state
        ref.compute_dt()
        ref.step_device()
    U_start = ref.U.clone()
    ref.compute_dt()
    dt0 = ref.dt_buf[0].clone()
    ref.step_device()
    want_state, want_dt = ref.U.clone(), float(ref.dt_buf[0].item())

    s = Solver(cfg)
    L, g = s.L, cfg.grid
    U = U_start.clone()
    U1, U2 = torch.zeros_like(U), torch.zeros_like(U)
    F = torch.zeros((28, g.nxg, s.ld), dtype=torch.float64, device=U.device)
    dt = torch.zeros(2, dtype=torch.float64, device=U.device)
    dt[0] = dt0
    st = torch.cuda.current_stream().cuda_stream
    p6 = lambda t: _lib.ptrs6([t[m].data_ptr() for m in range(6)])
    pF = _lib.ptrs28([F[m].data_ptr() for m in range(28)])
    stages = ((U, U1, 0.0, 1.0), (U1, U2, 0.75, 0.25), (U2, U, 1.0 / 3.0, 2.0 / 3.0))
    for cur, dst, a, b in stages:
        _lib.check(L.wamp_compute_fluxes(s.ctx, p6(cur), pF, st), "wamp_compute_fluxes")
        _lib.check(L.wamp_limit_update(s.ctx, p6(U), p6(cur), pF, p6(dst), a, b, dt.data_ptr(),
                                       None, st), "wamp_limit_update")
    _lib.check(L.wamp_get_dt(s.ctx, p6(U), dt.data_ptr() + 8, st), "wamp_get_dt")
    torch.cuda.synchronize()
    si, sj = g.interior()
    got, want = U.cpu().numpy(), want_state.cpu().numpy()
    for m in range(6):
        assert np.array_equal(got[m][si, sj], want[m][si, sj]), m
    assert float(dt[1].item()) == want_dt
    s.close()
    ref.close()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("scheme", ["sc", "m10", "mp5"])
def test_split_calls_equal_one_step_3d(gpu_ok, scheme, precision):
    """wamp3_compute_fluxes into 48 caller-owned FP64 planes, then
    wamp3_limit_update per RK stage, then wamp3_get_dt, reproduce wamp3_step
    bit for bit (in mixed precision wamp3_step keeps 42 of the planes in FP32)."""
    import dataclasses

    import torch

    from wamp import _lib, threed
    from wamp.config import SCHEMES

    cfg, U0 = threed.shock_sphere(12, scheme=SCHEMES[scheme])
    cfg = dataclasses.replace(cfg, precision=precision)
    ref = threed.Solver3(cfg)
    ref.set_state(U0)
    ref.prepare()
    for _ in range(10):  # the shock on the sphere
        ref.compute_dt()
        ref.step_device()
    U_start = ref.U.clone()
    ref.compute_dt()
    dt0 = ref.dt_buf[0].clone()
    ref.step_device()
    want_state, want_dt = ref.U.clone(), float(ref.dt_buf[0].item())

    s = threed.Solver3(cfg)
    L, g = s.L, cfg.grid
    U = U_start.clone()
    U1, U2 = torch.zeros_like(U), torch.zeros_like(U)
    F = torch.zeros((48,) + tuple(U.shape[1:]), dtype=torch.float64, device=U.device)
    dt = torch.zeros(2, dtype=torch.float64, device=U.device)
    dt[0] = dt0
    st = torch.cuda.current_stream().cuda_stream
    p7 = lambda t: _lib.PtrArr7(*[t[m].data_ptr() for m in range(7)])
    pF = _lib.ptrs48([F[m].data_ptr() for m in range(48)])
    stages = ((U, U1, 0.0, 1.0), (U1, U2, 0.75, 0.25), (U2, U, 1.0 / 3.0, 2.0 / 3.0))
    for cur, dst, a, b in stages:
        _lib.check(L.wamp3_compute_fluxes(s.ctx, p7(cur), pF, st), "wamp3_compute_fluxes")
        _lib.check(L.wamp3_limit_update(s.ctx, p7(U), p7(cur), pF, p7(dst), a, b, dt.data_ptr(),
                                        None, st), "wamp3_limit_update")
    _lib.check(L.wamp3_get_dt(s.ctx, p7(U), dt.data_ptr() + 8, st), "wamp3_get_dt")
    torch.cuda.synchronize()
    si, sj, sk = g.interior()
    got, want = U.cpu().numpy(), want_state.cpu().numpy()
    moved = np.abs(want - U_start.cpu().numpy())[:, si, sj, sk].max()
    assert moved > 0
    for m in range(7):
        assert np.array_equal(got[m][si, sj, sk], want[m][si, sj, sk]), m
    assert float(dt[1].item()) == want_dt
    s.close()
    ref.close()


def test_split_3d_rejects_bad_plane_lists(gpu_ok):
    """Face planes that are not equally spaced, or that overlap, and an output
    state aliasing the input are refused before anything is launched."""
    import torch

    from wamp import _lib, threed

    cfg, U0 = threed.shock_sphere(4)
    s = threed.Solver3(cfg)
    L = s.L
    U = torch.zeros_like(s.U)
    F = torch.zeros((49,) + tuple(U.shape[1:]), dtype=torch.float64, device=U.device)
    st = torch.cuda.current_stream().cuda_stream
    p7 = lambda t: _lib.PtrArr7(*[t[m].data_ptr() for m in range(7)])
    gap = [F[m if m < 20 else m + 1].data_ptr() for m in range(48)]
    assert L.wamp3_compute_fluxes(s.ctx, p7(U), _lib.ptrs48(gap), st) == -2
    assert b"equally spaced" in L.wamp_last_error()
    base = F[0].data_ptr()
    overlap = [base + 8 * m for m in range(48)]
    assert L.wamp3_compute_fluxes(s.ctx, p7(U), _lib.ptrs48(overlap), st) == -2
    assert b"overlap" in L.wamp_last_error()
    pF = _lib.ptrs48([F[m].data_ptr() for m in range(48)])
    dt = torch.zeros(1, dtype=torch.float64, device=U.device)
    rc = L.wamp3_limit_update(s.ctx, p7(U), p7(U), pF, p7(U), 0.0, 1.0, dt.data_ptr(), None, st)
    assert rc == -2 and b"aliases" in L.wamp_last_error()
    s.close()
