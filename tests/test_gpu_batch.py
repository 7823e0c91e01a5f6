"""Batched solves: every case is bitwise the same as solved alone, whatever the
tuning, with or without a captured graph; end times freeze a case exactly."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cases(precision):
    from wamp import cases
    from wamp.config import SCHEME_MP5

    rng = np.random.default_rng(11)
    out = [cases.shock_cylinder(96, 48, scheme=SCHEME_MP5),
           cases.triple_point(70, 30, scheme=SCHEME_MP5),
           cases.batch_case(rng, scheme=SCHEME_MP5, grids=[(80, 64)]),
           cases.batch_case(rng, scheme=SCHEME_MP5, grids=[(130, 40)])]
    return [(dataclasses.replace(c, precision=precision), U) for c, U in out]


def _alone(cfg, U0, steps):
    from wamp.solver import Solver

    s = Solver(cfg)
    s.set_state(U0)
    s.prepare()
    s.run(steps)
    out = s.state()
    s.close()
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_equals_alone_under_all_tunings(gpu_ok, precision):
    from wamp.batch import Batch, Tuning

    cs = _cases(precision)
    steps = 6
    ref = [_alone(c, U, steps) for c, U in cs]
    # limiter blocks per CU 0: one wave per item; 1, 2, 4: persistent waves
    # striding over the items
    for tuning, graph in ((Tuning(0, 3, 2, 16), 0), (Tuning(1, 2, 4, 8), 3), (Tuning(0, 1, 1, 32), 6),
                          (Tuning(1, 3, 2, 26), 2), (Tuning(0, 3, 0, 20), 0), (Tuning(1, 3, 0, 26), 2),
                          (Tuning(0, 3, 0, 5), 0), (Tuning(1, 2, 2, 100), 2),
                          (Tuning(0, 3, 0, 0), 0), (Tuning(1, 3, 2, 0), 3)):
        b = Batch(cs[0][0].scheme, precision, max_cases=6)
        for k, (c, U) in enumerate(cs):
            b.set_case(k + 1, c, U)  # slot 0 and 5 stay empty
        b.set_tuning(tuning)
        b.commit()
        rows = b.work()[2]
        assert rows == tuning.lu_seg if tuning.lu_seg else 4 <= rows <= 4096
        b.prepare()
        if graph:
            b.capture(graph)
        done = 0
        while done < steps:
            done += b.step()
        assert done == steps
        for k, (c, _) in enumerate(cs):
            si, sj = c.grid.interior()
            got = b.state(k + 1)
            assert np.array_equal(got[:, si, sj], ref[k][:, si, sj]), (tuning, graph, k)
        b.close()


def test_end_time_is_hit_exactly_and_freezes_the_case(gpu_ok):
    from wamp.batch import Batch

    cs = _cases(1)
    b = Batch(cs[0][0].scheme, 1, max_cases=4)
    t_end = [3e-3, 0.0, 1e-3, 2e-3]
    for k, (c, U) in enumerate(cs):
        b.set_case(k, c, U, t0=0.0, t_end=t_end[k])
    b.commit()
    b.prepare()
    b.capture(4)
    for _ in range(100):
        b.step()
        t, dt, done = b.status()
        if done[0] and done[2] and done[3]:
            break
    assert done[0] and not done[1] and done[2] and done[3]
    for k in (0, 2, 3):
        assert t[k] == pytest.approx(t_end[k], rel=1e-12)
    frozen = b.state(2)
    b.step()
    assert np.array_equal(b.state(2), frozen)
    b.close()


def test_autotune_keeps_results(gpu_ok):
    from wamp.batch import Batch, Tuning

    cs = _cases(1)
    ref = [_alone(c, U, 4) for c, U in cs]
    b = Batch(cs[0][0].scheme, 1, max_cases=len(cs))
    for k, (c, U) in enumerate(cs):
        b.set_case(k, c, U)
    b.commit()
    b.prepare()
    best = b.autotune(steps=1, space=[Tuning(0, 3, 2, 16), Tuning(1, 2, 2, 8)])
    assert best in (Tuning(0, 3, 2, 16), Tuning(1, 2, 2, 8))
    for _ in range(4):
        b.step()
    for k, (c, _) in enumerate(cs):
        si, sj = c.grid.interior()
        assert np.array_equal(b.state(k)[:, si, sj], ref[k][:, si, sj])
    # it restores states and start times, so it refuses once the cases moved
    with pytest.raises(RuntimeError):
        b.autotune(steps=1, space=[Tuning(0, 3, 2, 16)])
    b.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_with_degenerate_cases_equals_alone(gpu_ok, precision):
    """One-cell, one-row and ragged cases beside a normal one in the same
    persistent launches: each bitwise equal to its solve alone."""
    from wamp import cases
    from wamp.batch import Batch, Tuning
    from wamp.config import SCHEME_M10

    shapes = [(1, 1), (6, 1), (3, 2), (130, 3), (9, 61), (96, 48)]
    cs = [(dataclasses.replace(c, precision=precision), U)
          for c, U in (cases.shock_cylinder(nx, ny, scheme=SCHEME_M10) for nx, ny in shapes)]
    steps = 5
    ref = [_alone(c, U, steps) for c, U in cs]
    for tuning in (Tuning(0, 3, 2, 8), Tuning(1, 2, 2, 26)):
        b = Batch(SCHEME_M10, precision, max_cases=len(cs))
        for k, (c, U) in enumerate(cs):
            b.set_case(k, c, U)
        b.set_tuning(tuning)
        b.commit()
        b.prepare()
        done = 0
        while done < steps:
            done += b.step()
        for k, (c, _) in enumerate(cs):
            si, sj = c.grid.interior()
            assert np.array_equal(b.state(k)[:, si, sj], ref[k][:, si, sj]), (tuning, shapes[k])
        b.close()


def test_failed_set_case_keeps_the_slot_and_the_graph(gpu_ok):
    """A case whose scratch cannot be allocated (a grid far beyond HBM) is
    refused; the slot keeps its case, the batch stays committed and its
    captured graph keeps stepping it, bitwise as the case alone; later
    launches elsewhere do not see the refused allocation's error."""
    import ctypes as C

    from wamp import _lib
    from wamp.batch import Batch

    cs = _cases(1)[:1]
    ref = _alone(cs[0][0], cs[0][1], 4)
    b = Batch(cs[0][0].scheme, 1, max_cases=1)
    b.set_case(0, *cs[0])
    b.commit()
    b.prepare()
    b.capture(2)
    b.step()
    U = b.states[0]
    huge = _lib.WampGrid(60000, 60000, 5, 60010, 1e-4, 1e-4, (C.c_int32 * 4)(0, 0, 1, 1))
    eos = _lib.WampEos(*cs[0][0].eos.as_tuple())
    rc = b.L.wamp_batch_set_case(b.h, 0, C.byref(huge), C.byref(eos),
                                 _lib.ptrs6([U[m].data_ptr() for m in range(6)]), 0.0, 0.0)
    assert rc > 0 and "hipMalloc" in b.L.wamp_last_error().decode()
    assert b.step() == 2        # still committed, the graph replays
    # the refused allocation is reported once, not again by later launches
    assert np.array_equal(_alone(cs[0][0], cs[0][1], 4), ref)
    si, sj = cs[0][0].grid.interior()
    assert np.array_equal(b.state(0)[:, si, sj], ref[:, si, sj])
    b.close()
