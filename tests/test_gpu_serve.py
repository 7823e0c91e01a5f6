"""Serving: cases admitted into and retired from a running batch give the same
result as each case solved alone by the plain solver, bitwise, in both
precisions; admission respects the HBM budget and slots are reused."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class RoundClock:
    """Virtual clock in units of served rounds (deterministic admission)."""

    def __init__(self):
        self.srv = None
        self.idle = 0.0

    def __call__(self):
        return float(self.srv.rounds) + self.idle

    def sleep(self, d):
        self.idle += d


def _requests(precision):
    from wamp import cases, serve
    from wamp.config import SCHEME_MP5

    rng = np.random.default_rng(21)

    def req(rid, arrival, cfg_U, t_end):
        cfg, U = cfg_U
        return serve.Request(arrival, rid, dataclasses.replace(cfg, precision=precision), U, t_end)

    return [
        req(0, 0.0, cases.shock_cylinder(96, 48, scheme=SCHEME_MP5), 0.05),
        req(1, 0.0, cases.triple_point(70, 30, scheme=SCHEME_MP5), 0.5),
        req(2, 3.0, cases.batch_case(rng, scheme=SCHEME_MP5, grids=[(80, 64)]), 0.03),
        req(3, 5.0, cases.batch_case(rng, scheme=SCHEME_MP5, grids=[(130, 40)]), 0.04),
        req(4, 6.0, cases.triple_point(64, 32, scheme=SCHEME_MP5), 0.3),
        # arrives after everything else finished: the loop idles, then serves it;
        # its initial state is built in device memory at admission
        serve.Request(1000.0, 5, dataclasses.replace(cases.shock_cylinder(64, 32, scheme=SCHEME_MP5)[0],
                                                     precision=precision), None, 0.02,
                      build=lambda: cases.shock_droplets_device(
                          cases.shock_cylinder(64, 32)[0].grid, cases.WATER_ND, 2.4, 0.8,
                          [(1.5, 1.0, 0.5, 0.0)])),
    ]


def _alone(req):
    from wamp.solver import Solver

    s = Solver(req.cfg)
    s.set_state(req.state0())
    s.prepare()
    s.advance_to(req.t_end, log=None)
    out, t = s.state(), s.time
    s.close()
    return out, t


@pytest.mark.parametrize("precision", [0, 1])
def test_served_results_equal_alone(gpu_ok, precision):
    from wamp.serve import Server

    reqs = _requests(precision)
    # with requests 0-2 resident, request 3 does not fit and waits for memory
    budget = float(sum(r.nbytes for r in reqs[:4]) - 1)
    clock = RoundClock()
    srv = Server(scheme=reqs[0].cfg.scheme, precision=precision, max_slots=4, hbm_budget=budget,
                 round_steps=2, clock=clock, sleep=clock.sleep)
    clock.srv = srv
    for r in reqs:
        srv.submit(r)
    srv.run()
    assert len(srv.results) == len(reqs)
    assert srv.peak_used <= budget
    assert srv.peak_resident >= 3
    by_id = {r.rid: r for r in srv.results}
    first_out = min(r.finished_round for r in srv.results)
    assert by_id[3].admitted_round == first_out > 5  # waited for memory

    # cases joined a batch that was already running
    r0 = by_id[0]
    assert any(by_id[k].admitted_round > r0.admitted_round and by_id[k].admitted_round < r0.finished_round
               for k in (2, 3, 4))
    assert by_id[5].admitted >= 1000.0

    for req in reqs:
        ref, t_ref = _alone(req)
        got = by_id[req.rid]
        assert got.t == t_ref == pytest.approx(req.t_end, rel=1e-12)
        si, sj = req.cfg.grid.interior()
        assert np.array_equal(got.state[:, si, sj], ref[:, si, sj]), req.rid
    srv.close()


def test_request_larger_than_budget_is_refused(gpu_ok):
    from wamp.serve import Server

    reqs = _requests(1)
    srv = Server(scheme=reqs[0].cfg.scheme, precision=1, max_slots=2,
                 hbm_budget=reqs[0].nbytes - 1)
    with pytest.raises(ValueError):
        srv.submit(reqs[0])
    srv.close()
