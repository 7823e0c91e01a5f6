"""Oracle-generated regression fixtures (tests/golden/make_golden.py): the oracle
must keep reproducing them; the HIP path must match them to the parity bound."""
import glob
import os

import numpy as np
import pytest

from conftest import rel_l1

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = sorted(glob.glob(os.path.join(HERE, "shock_cylinder_*.npz")))


def _load(path):
    d = np.load(path, allow_pickle=False)
    return d, str(d["scheme"]), int(d["nx"]), int(d["ny"]), int(d["steps"])


@pytest.mark.parametrize("path", FILES, ids=os.path.basename)
def test_oracle_reproduces_golden(path):
    from oracle.oracle import Oracle
    from wamp import cases
    from wamp.config import SCHEMES

    d, scheme, nx, ny, steps = _load(path)
    cfg, U0 = cases.shock_cylinder(nx, ny, scheme=SCHEMES[scheme])
    orc = Oracle(cfg)
    orc.set_state(U0)
    orc.prepare()
    dts = orc.run(steps)
    si, sj = cfg.grid.interior()
    S = orc.state()[:, si, sj]
    assert np.allclose(dts, d["dts"], rtol=1e-14, atol=0)
    for m in range(6):
        assert rel_l1(S[m], d["state"][m]) <= 1e-14, m


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=os.path.basename)
def test_hip_matches_golden(gpu_ok, path):
    from wamp import cases
    from wamp.config import SCHEMES
    from wamp.solver import Solver

    d, scheme, nx, ny, steps = _load(path)
    cfg, U0 = cases.shock_cylinder(nx, ny, scheme=SCHEMES[scheme])
    s = Solver(cfg)
    s.set_state(U0)
    s.prepare()
    dts = s.run(steps)
    si, sj = cfg.grid.interior()
    S = s.state()[:, si, sj]
    assert np.allclose(dts, d["dts"], rtol=1e-13, atol=0)
    for m in range(6):
        assert rel_l1(S[m], d["state"][m]) <= 1e-12, m
