"""Mixed precision (FP32 faces, FP64 pressure recovery / limiter / RK) against FP64."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-4  # relative L1 versus the FP64 solution


@pytest.mark.parametrize("scheme", ["sc", "m10", "mp5"])
def test_mixed_close_to_fp64_oracle(gpu_ok, scheme):
    from oracle.oracle import Oracle
    from wamp import accuracy, cases
    from wamp.config import PREC_MIXED, SCHEMES
    from wamp.solver import Solver

    cfg, U0 = cases.shock_cylinder(256, 128, scheme=SCHEMES[scheme])
    orc = Oracle(cfg)
    orc.set_state(U0)
    orc.prepare()
    orc.run(200)
    ref = orc.state()
    s = Solver(dataclasses.replace(cfg, precision=PREC_MIXED))
    s.set_state(U0)
    s.prepare()
    s.run(200)
    err = accuracy.compare(cfg, s.state(), ref)
    print(scheme, err)
    assert err["finite"]
    for k in ("rho", "p", "alpha1"):
        assert err[k] <= TOL, (k, err)
    assert err["centroid_shift_cells"] < 1e-3


def test_mixed_is_not_fp64(gpu_ok):
    """The mixed path really runs different (FP32) face arithmetic."""
    from wamp import cases
    from wamp.config import PREC_MIXED, SCHEME_MP5
    from wamp.solver import Solver

    cfg, U0 = cases.shock_cylinder(64, 32, scheme=SCHEME_MP5)
    out = []
    for prec in (0, PREC_MIXED):
        s = Solver(dataclasses.replace(cfg, precision=prec))
        s.set_state(U0)
        s.prepare()
        s.run(5)
        out.append(s.state())
    d = np.abs(out[0] - out[1]).max()
    assert 0.0 < d < 1e-3


def test_scratch_holds_fp32_face_planes_in_mixed_precision(gpu_ok):
    """FP64: 28 face + 12 stage planes; mixed: 4 FP64 energy-flux planes and 24
    FP32 face planes + 12 stage planes (plus the dt accumulator)."""
    from wamp import cases
    from wamp.config import PREC_FP64, PREC_MIXED, SCHEME_MP5
    from wamp.solver import Solver

    cfg, _ = cases.shock_cylinder(160, 80, scheme=SCHEME_MP5)
    got = {}
    for prec in (PREC_FP64, PREC_MIXED):
        s = Solver(dataclasses.replace(cfg, precision=prec))
        plane = cfg.grid.nxg * s.ld
        got[prec] = (s.scratch_bytes(), plane)
        s.close()
    b64, plane = got[PREC_FP64]
    bmix, _ = got[PREC_MIXED]
    assert b64 == (40 * plane + 32) * 8
    assert bmix == 4 * plane * 8 + 24 * plane * 4 + (12 * plane + 32) * 8
