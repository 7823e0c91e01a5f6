"""The oracle's operation-counting build (oracle/opcount.h, liborc_count.so).

Every double of the oracle becomes a counting type that performs the same
IEEE operation, so the counting build must step bit for bit like the plain
build. Two of its counts have closed forms fixed by the reference's structure:
- pow: each face evaluates the entropy p / rho^gamma of its 8 stencil cells
  i-3..i+4 (SC:1019-1026, cell_entropy SC:309-314), so there are 8 pows per
  face and stage.
- sqrt: each face evaluates 9 sound speeds (sound_speed SC:98-103): one for the averaged
  state of the characteristic projection, and both sides of hllc and of
  hllc_src, for the high- and the first-order flux.
With (nx + 1) ny x-faces and nx (ny + 1) y-faces per stage, one SSP-RK3 step
has 24 pows and 27 sqrts per face.
"""
import numpy as np
import pytest

from wamp import cases
from wamp.config import SCHEMES


def _faces(g):
    return (g.nx + 1) * g.ny + g.nx * (g.ny + 1)


@pytest.mark.parametrize("scheme", ["sc", "m10", "mp5"])
def test_counting_build_steps_bitwise_like_the_plain_build(scheme):
    from oracle.oracle import Oracle

    cfg, U0 = cases.shock_cylinder(48, 24, scheme=SCHEMES[scheme])
    out = []
    for counting in (False, True):
        o = Oracle(cfg, threads=1, counting=counting)
        o.set_state(U0)
        o.prepare()
        dts = o.run(3)
        out.append((o.state(), dts))
        o.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("scheme", ["sc", "m10", "mp5"])
@pytest.mark.parametrize("nx, ny", [(32, 16), (37, 23)])
def test_pow_and_sqrt_counts_follow_the_face_structure(scheme, nx, ny):
    from oracle.oracle import Oracle

    cfg, U0 = cases.shock_cylinder(nx, ny, scheme=SCHEMES[scheme])
    o = Oracle(cfg, counting=True)
    o.set_state(U0)
    o.prepare()
    dt = o.get_dt()
    o.opcounts(reset=True)
    o.step(dt)
    c = o.opcounts(reset=True)
    o.close()
    f = _faces(cfg.grid)
    assert c["pow"] == 24 * f
    assert c["sqrt"] == 27 * f
    for k in ("add", "mul", "div", "cmp", "fabs"):
        assert c[k] > 0, k


def test_opcounts_need_the_counting_build():
    from oracle.oracle import Oracle

    cfg, U0 = cases.shock_cylinder(16, 8, scheme=SCHEMES["sc"])
    o = Oracle(cfg, threads=1)
    with pytest.raises(RuntimeError):
        o.opcounts()
    o.close()
