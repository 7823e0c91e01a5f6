"""Three-dimensional extension (wamp3_*) against the 2D path.

The reference has no 3D solver, so the 3D kernels are pinned through exact
reductions to the 2D kernels, which the oracle pins: a z-invariant state
with w = 0 must take exactly the 2D step of its (x, y) slice, and a
y-invariant state with v = 0 exactly the 2D step of its (x, z) slice (2D y
as z, 2D v as w), for the same dt sequence.  Every sweep, the z-sweep
included, is covered by one of the two embeddings; every scheme the 3D
library builds (SC, M10, MP5) is checked.  A genuinely 3D case (shock on a
water sphere) then checks finiteness and the y <-> z symmetry of the setup.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(scheme, embed, steps, nx=64, ny=32, thick=4):
    from wamp import cases, threed
    from wamp.config import SCHEMES
    from wamp.solver import Solver

    cfg2, U2 = cases.shock_cylinder(nx, ny, scheme=SCHEMES[scheme])
    cfg3, U3 = (threed.from_2d_xy if embed == "xy" else threed.from_2d_xz)(cfg2, U2, thick)
    s2 = Solver(cfg2)
    s2.set_state(U2)
    s2.prepare()
    s3 = threed.Solver3(cfg3)
    s3.set_state(U3)
    s3.prepare()
    s2.compute_dt()
    for _ in range(steps):
        dt = float(s2.dt_buf[0].item())
        s2.step_device()
        s3.step_with(dt)
    out2, out3 = s2.state(), s3.state()
    s2.close()
    s3.close()
    return cfg2, out2, cfg3, out3


def _slices(cfg2, U2, cfg3, U3, embed):
    """Yield (name, 3D slice, 2D field) over every interior slice of the embedding."""
    from wamp.threed import XY_MAP, XZ_MAP

    si, sj = cfg2.grid.interior()
    g3 = cfg3.grid
    if embed == "xy":
        for k in range(g3.g, g3.g + g3.nz):
            for m2, m3 in enumerate(XY_MAP):
                yield f"k={k} m={m3}", U3[m3][si, sj, k], U2[m2][si, sj]
            yield f"k={k} rw", U3[4][si, sj, k], np.zeros_like(U2[0][si, sj])
    else:
        for j in range(g3.g, g3.g + g3.ny):
            for m2, m3 in enumerate(XZ_MAP):
                yield f"j={j} m={m3}", U3[m3][si, j, sj], U2[m2][si, sj]
            yield f"j={j} rv", U3[3][si, j, sj], np.zeros_like(U2[0][si, sj])


@pytest.mark.parametrize("embed", ["xy", "xz"])
@pytest.mark.parametrize("scheme", ["sc", "m10", "mp5"])
def test_invariant_3d_state_steps_exactly_as_2d(gpu_ok, scheme, embed):
    cfg2, U2, cfg3, U3 = _pair(scheme, embed, steps=10)
    assert np.isfinite(U3).all()
    for name, a, b in _slices(cfg2, U2, cfg3, U3, embed):
        if not np.array_equal(a, b):
            d = np.abs(a - b)
            k = np.unravel_index(np.argmax(d), d.shape)
            pytest.fail(f"{scheme} {embed} {name}: {np.count_nonzero(a != b)} cells differ, "
                        f"max |diff| {d[k]:.3e} at {k} ({a[k]!r} vs {b[k]!r})")


@pytest.mark.parametrize("embed", ["xy", "xz"])
def test_invariant_3d_state_against_oracle(gpu_ok, embed):
    """The embedded 2D shock-cylinder case (MP5), 10 steps of the 3D FP64
    solver with the oracle's dt sequence, against the oracle's 2D state on
    every interior slice."""
    from conftest import rel_l1
    from oracle.oracle import Oracle
    from wamp import cases, threed
    from wamp.config import SCHEME_MP5

    cfg2, U2 = cases.shock_cylinder(64, 32, scheme=SCHEME_MP5)
    cfg3, U3 = (threed.from_2d_xy if embed == "xy" else threed.from_2d_xz)(cfg2, U2, 4)
    orc = Oracle(cfg2)
    orc.set_state(U2)
    orc.prepare()
    s3 = threed.Solver3(cfg3)
    s3.set_state(U3)
    s3.prepare()
    for _ in range(10):
        dt = orc.get_dt()
        orc.step(dt)
        s3.step_with(dt)
    ref, got = orc.state(), s3.state()
    orc.close()
    s3.close()
    worst = 0.0
    for name, a, b in _slices(cfg2, ref, cfg3, got, embed):
        if np.abs(b).sum() > 0:
            worst = max(worst, rel_l1(a, b))
        else:
            assert np.all(a == 0.0), name
    print(f"3D {embed} embedding vs oracle after 10 steps: max rel L1 {worst:.2e}")
    assert worst <= 1e-12


def test_one_step_reduction_on_a_ragged_grid(gpu_ok):
    """Grid sizes that are not multiples of the launch tiles (37 x 23)."""
    cfg2, U2, cfg3, U3 = _pair("mp5", "xz", steps=1, nx=37, ny=23)
    for name, a, b in _slices(cfg2, U2, cfg3, U3, "xz"):
        assert np.array_equal(a, b), name


@pytest.mark.parametrize("embed", ["xy", "xz"])
def test_reduction_holds_for_a_single_cell_thick_slab(gpu_ok, embed):
    """nz = 1 (or ny = 1): every neighbour along the invariant axis is a ghost."""
    cfg2, U2, cfg3, U3 = _pair("m10", embed, steps=5, thick=1)
    for name, a, b in _slices(cfg2, U2, cfg3, U3, embed):
        assert np.array_equal(a, b), name


def test_padded_z_stride_gives_the_same_result(gpu_ok):
    from wamp import threed

    cfg, U0 = threed.shock_sphere(12)
    out = []
    for ldz in (None, cfg.grid.nzg + 5):
        s = threed.Solver3(cfg, ldz=ldz)
        s.set_state(U0)
        s.prepare()
        s.run(5)
        out.append(s.state())
        s.close()
    assert np.array_equal(out[0], out[1])


def test_shock_sphere_is_finite_and_keeps_its_y_z_symmetry(gpu_ok):
    """The sphere sits on the y = z diagonal of a cube cross-section with y and z
    slip walls, so the flow is symmetric under (y, v) <-> (z, w).  The y- and
    z-sweeps are the same arithmetic; only the order in which their terms
    enter the residual sums differs, so the symmetry holds to round-off."""
    from wamp import threed

    cfg, U0 = threed.shock_sphere(24)
    s = threed.Solver3(cfg)
    s.set_state(U0)
    s.prepare()
    dts = s.run(30)
    U = s.state()
    s.close()
    assert np.all(dts > 0) and np.isfinite(U).all()
    si, sj, sk = cfg.grid.interior()
    I = U[:, si, sj, sk]
    T = np.swapaxes(I, 2, 3)
    swap = [0, 1, 2, 4, 3, 5, 6]  # v <-> w
    for m in range(7):
        scale = np.abs(I[m]).max()
        err = np.abs(I[m] - T[swap[m]]).max() / scale
        assert err < 1e-11, (m, err)
    # the shock moved into the sphere: alpha1 and the x-momentum changed
    assert np.abs(I[2] - U0[2][si, sj, sk]).max() > 0.1 * np.abs(U0[2]).max()


def test_scratch_is_68_planes(gpu_ok):
    from wamp import threed

    cfg, U0 = threed.shock_sphere(8)
    s = threed.Solver3(cfg)
    g = cfg.grid
    assert s.scratch_bytes() == (68 * g.nxg * g.nyg * g.nzg + 32) * 8
    s.close()


@pytest.mark.parametrize("scheme", ["sc", "m10", "mp5"])
def test_mixed_precision_stays_close_to_fp64(gpu_ok, scheme):
    """FP32-mixed 3D faces (energy flux, pressure, indicators, limiters FP64)
    against the FP64 3D step on the shock sphere, same dt sequence: within the
    1e-4 relative L1 of the 2D mixed path, and really different arithmetic."""
    import dataclasses

    from wamp import threed
    from wamp.config import SCHEMES

    cfg, U0 = threed.shock_sphere(24, scheme=SCHEMES[scheme])
    runs = []
    for prec in (0, 1):
        s = threed.Solver3(dataclasses.replace(cfg, precision=prec))
        s.set_state(U0)
        s.prepare()
        runs.append(s)
    runs[0].compute_dt()
    for _ in range(40):
        dt = float(runs[0].dt_buf[0].item())
        runs[0].step_device()
        runs[1].step_with(dt)
    ref, mixed = runs[0].state(), runs[1].state()
    for s in runs:
        s.close()
    err = threed.compare3(cfg, mixed, ref)
    print(f"3D mixed vs fp64 after 40 steps ({scheme}):", err)
    assert err["finite"]
    for k in ("rho", "p", "alpha1"):
        assert err[k] <= 1e-4, (k, err)
    assert 0.0 < max(err["rho"], err["p"])


@pytest.mark.parametrize("scheme", ["sc", "m10", "mp5"])
@pytest.mark.parametrize("embed", ["xy", "xz"])
def test_mixed_reduction_is_close_to_the_2d_mixed_step(gpu_ok, embed, scheme):
    """Mixed 3D against mixed 2D on the embedded case.  Not bitwise: the 2D
    mixed face kernel recomputes pressure and entropy in the mixed unit (one-
    ulp division, exp(g log rho) entropy), while the 3D mixed faces read the
    FP64 unit's pressure and entropy planes; a one-ulp FP64 difference can
    flip an FP32 rounding.  Measured after 10 steps on one MI355X: 4.2e-9 (SC),
    1.7e-8 (M10), 4.2e-9 (MP5), the same for both embeddings; the
    bound is 1e-6, a hundredth of the mixed path's 1e-4 tolerance."""
    import dataclasses

    from wamp import cases, threed
    from wamp.config import PREC_MIXED, SCHEMES
    from wamp.solver import Solver

    cfg2, U2 = cases.shock_cylinder(64, 32, scheme=SCHEMES[scheme])
    cfg2 = dataclasses.replace(cfg2, precision=PREC_MIXED)
    cfg3, U3 = (threed.from_2d_xy if embed == "xy" else threed.from_2d_xz)(cfg2, U2)
    cfg3 = dataclasses.replace(cfg3, precision=PREC_MIXED)
    s2 = Solver(cfg2)
    s2.set_state(U2)
    s2.prepare()
    s3 = threed.Solver3(cfg3)
    s3.set_state(U3)
    s3.prepare()
    s2.compute_dt()
    for _ in range(10):
        dt = float(s2.dt_buf[0].item())
        s2.step_device()
        s3.step_with(dt)
    out2, out3 = s2.state(), s3.state()
    s2.close()
    s3.close()
    worst = 0.0
    for name, a, b in _slices(cfg2, out2, cfg3, out3, embed):
        den = np.abs(b).sum()
        worst = max(worst, float(np.abs(a - b).sum() / den) if den > 0 else float(np.abs(a).max()))
    print(f"mixed {embed} {scheme} reduction: worst relative L1 {worst:.3e}")
    assert worst < 1e-6


def test_mixed_scratch_holds_fp32_face_planes(gpu_ok):
    import dataclasses

    from wamp import threed

    cfg, U0 = threed.shock_sphere(8)
    s = threed.Solver3(dataclasses.replace(cfg, precision=1))
    g = cfg.grid
    n = g.nxg * g.nyg * g.nzg
    assert s.scratch_bytes() == (6 + 20) * n * 8 + 42 * n * 4 + 32 * 8
    s.close()


def _generic_cube(n=20):
    """A smooth state varying in x, y and z with all three velocities non-zero,
    water-air mixture (alpha in [0.1, 0.9]), on an n^3 cube with slip walls."""
    from wamp import threed

    grid = threed.Grid3(n, n, n, bc=(1, 1, 1, 1, 1, 1))
    cfg = threed.Case3Config(grid, name="generic_cube")
    x = grid.centres(0)[:, None, None]
    y = grid.centres(1)[None, :, None]
    z = grid.centres(2)[None, None, :]
    tp = 2 * np.pi
    a1 = 0.5 + 0.4 * np.sin(tp * x) * np.cos(tp * y) * np.sin(tp * z + 0.3)
    p = 1.0 + 0.3 * np.cos(tp * (x + 0.5 * y + 0.25 * z))
    u = 0.2 * np.sin(tp * y + 0.1) * np.cos(tp * z)
    v = 0.15 * np.cos(tp * x) * np.sin(tp * z + 0.2)
    w = 0.1 * np.sin(tp * x + 0.4) * np.cos(tp * y)
    U = threed.conserved3(cfg.eos, a1, np.full_like(a1, 1000.0), np.full_like(a1, 1.0), u, v, w, p)
    out = np.zeros_like(U)
    si, sj, sk = grid.interior()
    out[:, si, sj, sk] = U[:, si, sj, sk]
    return cfg, out


def _steps(cfg, U, dt, n):
    from wamp import threed

    s = threed.Solver3(cfg)
    s.set_state(U)
    s.prepare()
    for _ in range(n):
        s.step_with(dt)
    out = s.state()
    s.close()
    return out


@pytest.mark.parametrize("axes", ["xy", "yz"])
@pytest.mark.parametrize("scheme", ["m10", "mp5"])
def test_generic_3d_state_commutes_with_axis_swaps(gpu_ok, axes, scheme):
    """Swapping two axes (and the two velocities) of a generic 3D state and
    stepping gives the swapped stepped state, to round-off.  x <-> y pins the
    x-sweep against the y-sweep (written with the 2D y-sweep's sign
    convention), y <-> z the y-sweep against the z-sweep, both with all three
    velocities and variation in all three directions."""
    import dataclasses

    from wamp import threed
    from wamp.config import SCHEMES

    cfg, U0 = _generic_cube()
    cfg = dataclasses.replace(cfg, scheme=SCHEMES[scheme])
    s = threed.Solver3(cfg)
    s.set_state(U0)
    s.prepare()
    dt = s.get_dt()
    s.close()
    if axes == "xy":
        ax, perm = (1, 2), [0, 1, 3, 2, 4, 5, 6]
    else:
        ax, perm = (2, 3), [0, 1, 2, 4, 3, 5, 6]
    swap = lambda U: np.swapaxes(U, *ax)[perm]
    A = _steps(cfg, U0, dt, 5)
    B = _steps(cfg, np.ascontiguousarray(swap(U0)), dt, 5)
    si, sj, sk = cfg.grid.interior()
    SA = swap(A)[:, si, sj, sk]
    B = B[:, si, sj, sk]
    moved = np.abs(A[:, si, sj, sk] - U0[:, si, sj, sk]).max(axis=(1, 2, 3))
    for m in range(7):
        scale = np.abs(B[m]).max()
        err = np.abs(SA[m] - B[m]).max() / scale
        assert err < 1e-12, (scheme, axes, m, err)
    assert (moved[2:5] > 0).all()  # every momentum changed over the 5 steps


def test_cli_writes_outputs_and_restarts(gpu_ok, tmp_path):
    """python -m wamp.threed: 4 steps saved every 2, then a restart from the
    step-2 file for 2 more steps reaches the step-4 state. The restart
    rebuilds E from p, so the match is to round-off, not bitwise."""
    from wamp import threed

    a, b = tmp_path / "a", tmp_path / "b"
    assert threed.main(["--n", "8", "--steps", "4", "--save-every", "2", "--out", str(a)]) == 0
    files = sorted(p.name for p in a.iterdir())
    assert files == ["shock_sphere_16x8x8_000002.npz", "shock_sphere_16x8x8_000004.npz"], files
    assert threed.main(["--n", "8", "--steps", "2", "--restart", str(a / files[0]), "--out", str(b)]) == 0
    assert [p.name for p in b.iterdir()] == [files[1]]
    with np.load(a / files[1]) as A, np.load(b / files[1]) as B:
        assert float(A["time"]) == pytest.approx(float(B["time"]), rel=1e-12)
        assert float(A["time"]) > 0
        for k in ("rho", "u", "v", "w", "p", "alpha1"):
            x, y = A[k], B[k]
            assert np.abs(x - y).sum() <= 1e-9 * np.abs(x).sum() + 1e-12, k


def _batch3_cases(scheme, precision):
    """Four uneven 3D cases of one scheme: a shock sphere (24x12x12), the
    generic cube (14^3), the shock cylinder embedded along z (40x20x3) and a
    one-cell-thick slab of it (40x1x20)."""
    import dataclasses

    from wamp import cases, threed
    from wamp.config import SCHEMES

    sch = SCHEMES[scheme]
    out = [threed.shock_sphere(6, scheme=sch)]
    cfg, U = _generic_cube(14)
    out.append((dataclasses.replace(cfg, scheme=sch), U))
    cfg2, U2 = cases.shock_cylinder(40, 20, scheme=sch)
    out.append(threed.from_2d_xy(cfg2, U2, 3))
    out.append(threed.from_2d_xz(cfg2, U2, 1))
    return [(dataclasses.replace(c, precision=precision), U) for c, U in out]


def _alone3(cfg, U0, steps):
    """Solver3 on one case: its own CFL dt sequence; returns (state, t)."""
    from wamp import threed

    s = threed.Solver3(cfg)
    s.set_state(U0)
    s.prepare()
    dts = s.run(steps)
    out = s.state()
    s.close()
    t = 0.0
    for dt in dts:
        t += float(dt)
    return out, t


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("scheme", ["sc", "m10", "mp5"])
def test_batch3_equals_each_case_alone(gpu_ok, scheme, precision):
    """Uneven 3D cases in one batch (one launch per kernel and stage over all
    of them, slots 0 and 5 empty), stepped one by one and by replays of a
    3-step graph: every case equals Solver3 run alone bit for bit, with the
    same time."""
    from wamp import threed
    from wamp.config import SCHEMES

    cs = _batch3_cases(scheme, precision)
    steps = 6
    ref = [_alone3(c, U, steps) for c, U in cs]
    for graph in (0, 3):
        b = threed.Batch3(SCHEMES[scheme], precision, max_cases=6)
        for k, (c, U) in enumerate(cs):
            b.set_case(k + 1, c, U)
        b.commit(graph_steps=graph)
        b.prepare()
        done = 0
        while done < steps:
            done += b.step()
        assert done == steps
        t, _, fin = b.status()
        for k, (c, _) in enumerate(cs):
            si, sj, sk = c.grid.interior()
            got, (want, t_want) = b.state(k + 1), ref[k]
            for m in range(7):
                assert np.array_equal(got[m][si, sj, sk], want[m][si, sj, sk]), (graph, k, m)
            assert t[k + 1] == t_want, (graph, k)
            assert not fin[k + 1]
        b.close()


def test_batch3_end_time_freezes_a_case(gpu_ok):
    """A case given an end time stops there (the last dt clipped, the case then
    frozen) while the others go on; a cleared slot can take a new case."""
    from wamp import threed
    from wamp.config import SCHEMES

    cs = _batch3_cases("m10", 0)
    (c0, U0), (c1, U1) = cs[0], cs[2]
    _, t2 = _alone3(c0, U0, 2)
    t_end = 1.5 * t2  # inside the third or fourth step
    b = threed.Batch3(SCHEMES["m10"], 0, max_cases=2)
    b.set_case(0, c0, U0, t_end=t_end)
    b.set_case(1, c1, U1)
    b.commit()
    b.prepare()
    for _ in range(8):
        b.step()
    t, dt, fin = b.status()
    assert fin[0] and abs(t[0] - t_end) <= 1e-14 * t_end and dt[0] == 0.0
    frozen = b.state(0)
    b.step()
    assert np.array_equal(b.state(0), frozen)
    want, t_want = _alone3(c1, U1, 9)
    si, sj, sk = c1.grid.interior()
    assert np.array_equal(b.state(1)[:, si, sj, sk], want[:, si, sj, sk])
    t, _, _ = b.status()
    assert t[1] == t_want
    b.clear_case(0)
    b.set_case(0, cs[1][0], cs[1][1])
    b.commit()
    b.prepare([0])
    b.step()
    want0, _ = _alone3(cs[1][0], cs[1][1], 1)
    si, sj, sk = cs[1][0].grid.interior()
    assert np.array_equal(b.state(0)[:, si, sj, sk], want0[:, si, sj, sk])
    b.close()
