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
