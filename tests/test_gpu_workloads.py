"""The solver's measured workloads, each in a GPU test against the oracle or
against the FP64 path the oracle pins:

  Mach-10 cylinder, M10    1024x512 for 2000 steps against the oracle build
                           that shares the solver's pow and tanh, bit for bit,
                           in eight checkpointed segments (the shock reaches
                           the cylinder near step 140 and has crossed it by
                           step 2000); 4096x2048 for 2000 steps (the shock meets
                           the cylinder near step 680) on the device with its
                           mirror symmetry and admissibility
  40 droplets, MP5         4096x2048: one FP64 step against the oracle; 1000
                           steps of mixed precision against FP64 (same dt
                           sequence) within 1e-4 in rho, p, alpha1
  64 uneven cases          seed 4, 500 steps: the FP64 batch equal bit for
                           bit to the plain solver on every case; the
                           smallest case against the oracle; the mixed batch
                           within 1e-4 of FP64 on cases 0-3
  served stream            a slice of the seed-5 request stream with its
                           largest request, under Poisson arrivals: every
                           result equal bit for bit to its solve alone; the
                           smallest requests in FP64 against the oracle
"""
import dataclasses
import os

import numpy as np
import pytest

from conftest import rel_l1

pytestmark = pytest.mark.gpu

THREADS = int(os.environ.get("OMP_NUM_THREADS", "16"))


def _interior_errs(cfg, got, ref):
    """Relative L1 per conserved field over the interior; the two momentum
    components relative to the L1 norm of the momentum (|rho u| + |rho v|): in
    these flows rho*v alone is a small field (zero before the shock meets a
    droplet off the axis), whose own L1 norm would magnify round-off."""
    si, sj = cfg.grid.interior()
    mom = float(np.abs(ref[2][si, sj]).sum() + np.abs(ref[3][si, sj]).sum())
    out = []
    for m in range(6):
        d = float(np.abs(got[m][si, sj] - ref[m][si, sj]).sum())
        den = mom if m in (2, 3) else float(np.abs(ref[m][si, sj]).sum())
        out.append(d / (den if den > 0 else 1.0))
    return out


def oracle_advance_to(orc, t_end):
    """The reference's time loop without saves (SC:2015-2059): dt from the state
    before each step, the last one clipped to t_end (SC:2025-2026)."""
    t, n = 0.0, 0
    dt = orc.get_dt()
    while t < t_end - 1.0e-14:
        if t + dt > t_end:
            dt = t_end - t
        orc.step(dt)
        t += dt
        n += 1
        dt = orc.get_dt()
    return t, n


# ---------------------------------------------------------------------------
# Mach-10 cylinder (M10 scheme)
# ---------------------------------------------------------------------------
LONG_STEPS, LONG_SEGMENTS = 2000, 8


class _LongRun:
    def __init__(self):
        from oracle.oracle import Oracle
        from wamp import cases
        from wamp.config import SCHEME_M10
        from wamp.solver import Solver

        self.cfg, U0 = cases.mach10_cylinder(1024, 512, scheme=SCHEME_M10)
        # the oracle build with the solver's pow and tanh: no elementary function
        # differs, so the comparison is not swamped by the flow's amplification
        # of last-place differences once the shock meets the water
        self.orc = Oracle(self.cfg, threads=THREADS, shared_math=True)
        self.orc.set_state(U0)
        self.orc.prepare()
        self.s = Solver(self.cfg)
        self.s.set_state(self.orc.state())
        self.done = 0
        self.t = 0.0


@pytest.fixture(scope="module")
def mach10_long(gpu_ok):
    run = _LongRun()
    yield run
    run.s.close()
    run.orc.close()


@pytest.mark.parametrize("segment", range(LONG_SEGMENTS))
def test_mach10_long_horizon_against_oracle(mach10_long, segment):
    """1024x512, 2000 steps in segments of 250 (each segment is a test, so a
    long run reports as it goes): every dt equal to the oracle's bit for bit,
    and the state equal to it at every checkpoint up to differences below
    1e-290 in absolute value (transverse momentum at its round-off floor).  (Against the oracle build with the C
    library's pow and tanh the run stays within the flow's own sensitivity:
    that oracle, started from its own state scaled by 1 +- 2^-52, drifts to
    1e-9 by step 150, right after the shock meets the cylinder.)"""
    run = mach10_long
    n = LONG_STEPS // LONG_SEGMENTS
    assert run.done == segment * n, "segments run in order"
    dts_ref = run.orc.run(n)
    dts = run.s.run(n)
    run.done += n
    run.t += float(dts_ref.sum())
    got, ref = run.s.state(), run.orc.state()
    errs = _interior_errs(run.cfg, got, ref)
    shock_x = 0.6 + 11.8 * run.t   # Mach 10 in air at rest: shock speed 11.83
    print(f"steps {run.done} (t {run.t:.4f}, shock near x={shock_x:.2f}): rel L1",
          ["%.1e" % e for e in errs])
    si, sj = run.cfg.grid.interior()
    a, b = got[:, si, sj], ref[:, si, sj]
    diff = a != b
    worst = float(np.abs(a - b).max())
    print(f"    cells differing per field {diff.sum(axis=(1, 2)).tolist()}, max |diff| {worst:.1e}, "
          f"max |value| there {float(np.abs(b[diff]).max()) if diff.any() else 0.0:.1e}")
    # equal except where both values sit at the round-off floor of an
    # (almost) vanishing field: differences below 1e-290 in absolute value,
    # from divisions whose quotient is subnormal
    assert worst <= 1e-290
    assert np.array_equal(dts, dts_ref)
    if segment == LONG_SEGMENTS - 1:
        assert shock_x > 1.25   # past the cylinder (centre 1.0, radius 0.25)


def test_mach10_full_size_2000_steps_symmetric_and_admissible(gpu_ok):
    """4096x2048 through the shock-cylinder interaction on the device."""
    from wamp import cases
    from wamp.config import SCHEME_M10
    from wamp.solver import Solver

    cfg, U0 = cases.mach10_cylinder(4096, 2048, scheme=SCHEME_M10)
    s = Solver(cfg)
    s.set_state(U0)
    del U0
    s.prepare()
    dts = s.run(2000)
    S = s.state()
    s.close()
    t = float(dts.sum())
    assert np.all(dts > 0) and np.all(np.isfinite(dts))
    si, sj = cfg.grid.interior()
    assert np.isfinite(S[:, si, sj]).all()
    ar1, ar2, al = S[0][si, sj], S[1][si, sj], S[5][si, sj]
    assert ar1.min() > 0.0 and ar2.min() > 0.0
    assert al.min() >= 1.0e-8 and al.max() <= 1.0 - 1.0e-8
    worst = 0.0
    for m in range(6):
        a = S[m][si, sj]
        b = a[:, ::-1] * (-1.0 if m == 3 else 1.0)
        den = np.abs(S[2][si, sj]).sum() if m == 3 else np.abs(a).sum()
        worst = max(worst, float(np.abs(a - b).sum() / den))
    print(f"4096x2048 after 2000 steps (t {t:.4f}, shock near x={0.6 + 11.8 * t:.2f}): "
          f"mirror asymmetry {worst:.2e}")
    assert 0.6 + 11.8 * t > 0.8   # the shock has hit the cylinder (front at x = 0.75)
    assert worst <= 1e-6


# ---------------------------------------------------------------------------
# 40 seeded droplets (MP5), 4096x2048
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def droplets40():
    from wamp import cases

    cfg, U0, drops = cases.multi_droplet(4096, 2048)
    assert len(drops) == 40 and sum(1 for d in drops if d[3] > 0) == 10
    return cfg, U0


def test_droplets_one_step_against_oracle(gpu_ok, droplets40):
    from oracle.oracle import Oracle
    from wamp.solver import Solver

    cfg, U0 = droplets40
    orc = Oracle(cfg, threads=THREADS)
    orc.set_state(U0)
    orc.prepare()
    U = orc.state()
    dt = orc.get_dt()
    orc.step(dt)
    ref = orc.state()
    orc.close()
    s = Solver(cfg)
    s.set_state(U)
    s.set_dt(dt)
    s.step_device()
    got = s.state()
    s.close()
    errs = _interior_errs(cfg, got, ref)
    print("40 droplets, 4096x2048, one FP64 step: rel L1", ["%.1e" % e for e in errs])
    assert max(errs) <= 1e-12


def test_droplets_mixed_1000_steps_against_fp64(gpu_ok, droplets40):
    """The mixed-precision workload: FP32 face arithmetic against the FP64 path
    (within 1e-9 of the oracle, above), same dt sequence, 1000 steps."""
    from wamp import accuracy

    cfg, U0 = droplets40
    rows = accuracy.paired_run(cfg, U0, 1000, checkpoints=(1, 100), variant="mixed")
    for r in rows:
        print(f"step {r['step']}: rho {r['rho']:.2e}  p {r['p']:.2e}  alpha1 {r['alpha1']:.2e}  "
              f"centroid shift {r['centroid_shift_cells']:.2e} cells")
        assert r["finite"]
        assert max(r["rho"], r["p"], r["alpha1"]) <= 1e-4
        assert r["centroid_shift_cells"] <= 1e-3
    assert rows[-1]["step"] == 1000


# ---------------------------------------------------------------------------
# 64 uneven cases (seed 4), 500 steps
# ---------------------------------------------------------------------------
BATCH_STEPS = 500


@pytest.fixture(scope="module")
def batch64():
    from wamp import cases
    from wamp.config import SCHEME_MP5

    rng = np.random.default_rng(4)
    return [cases.batch_case(rng, scheme=SCHEME_MP5) for _ in range(64)]


def _plain(cfg, U0, steps):
    from wamp.solver import Solver

    s = Solver(cfg)
    s.set_state(U0)
    s.prepare()
    s.run(steps)
    out = s.state()
    s.close()
    return out


def _batched(work, precision, steps, graph_steps=10):
    from wamp.batch import Batch, Tuning
    from wamp.config import SCHEME_MP5

    b = Batch(SCHEME_MP5, precision, max_cases=len(work))
    for k, (cfg, U0) in enumerate(work):
        b.set_case(k, dataclasses.replace(cfg, precision=precision), U0)
    b.set_tuning(Tuning(1, 3, 2, 26))
    b.commit()
    b.prepare()
    b.capture(graph_steps)
    done = 0
    while done < steps:
        done += b.step()
    assert done == steps
    return b


@pytest.fixture(scope="module")
def batch64_plain(gpu_ok, batch64):
    return [_plain(cfg, U0, BATCH_STEPS) for cfg, U0 in batch64]


def test_batch64_fp64_equals_plain_on_every_case(batch64, batch64_plain):
    grids = sorted(set((c.grid.nx, c.grid.ny) for c, _ in batch64))
    assert len(grids) == 5
    b = _batched(batch64, 0, BATCH_STEPS)
    for k, (cfg, _) in enumerate(batch64):
        si, sj = cfg.grid.interior()
        assert np.array_equal(b.state(k)[:, si, sj], batch64_plain[k][:, si, sj]), k
    b.close()


def test_batch64_smallest_case_against_oracle(batch64, batch64_plain):
    from oracle.oracle import Oracle

    k = min(range(len(batch64)), key=lambda q: batch64[q][0].grid.nx * batch64[q][0].grid.ny)
    cfg, U0 = batch64[k]
    orc = Oracle(cfg, threads=THREADS)
    orc.set_state(U0)
    orc.prepare()
    orc.run(BATCH_STEPS)
    errs = _interior_errs(cfg, batch64_plain[k], orc.state())
    orc.close()
    print(f"case {k} ({cfg.grid.nx}x{cfg.grid.ny}, {cfg.name}), {BATCH_STEPS} steps: rel L1",
          ["%.1e" % e for e in errs])
    assert max(errs) <= 1e-9


def test_batch64_mixed_spot_checks(batch64, batch64_plain):
    """Cases 0-3 in the mixed batch: each equal bit for bit to the plain mixed
    solver, and within 1e-4 of FP64 in rho, p and alpha1 -- or, where the flow
    itself amplifies round-off (a Mach 5-6 shock on a droplet rolls the
    interface up), within 50x of what one FP32 rounding of the initial state
    grows to in FP64 (the FP64 solver from the state scaled by 1 +- 2^-24 per
    cell).  With one dt sequence for both runs the mixed path measured 6-7x
    that floor on these cases."""
    from wamp import accuracy

    b = _batched(batch64, 1, BATCH_STEPS)
    for k in range(4):
        cfg, U0 = batch64[k]
        si, sj = cfg.grid.interior()
        assert np.array_equal(b.state(k)[:, si, sj],
                              _plain(dataclasses.replace(cfg, precision=1), U0, BATCH_STEPS)[:, si, sj])
        e = accuracy.compare(cfg, b.state(k), batch64_plain[k])
        floor = accuracy.compare(cfg, _plain(cfg, accuracy.perturbed(cfg, U0, 2.0 ** -24), BATCH_STEPS),
                                 batch64_plain[k])
        print(f"case {k} ({cfg.name}): mixed vs FP64 rho {e['rho']:.1e} p {e['p']:.1e} "
              f"alpha1 {e['alpha1']:.1e}; FP64 from a 2^-24 perturbation rho {floor['rho']:.1e} "
              f"p {floor['p']:.1e} alpha1 {floor['alpha1']:.1e}")
        assert e["finite"]
        for q in ("rho", "p", "alpha1"):
            assert e[q] <= max(1e-4, 50.0 * floor[q]), (k, q)
    b.close()


# ---------------------------------------------------------------------------
# The served request stream (seed 5)
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def stream():
    """The seed-5 request stream (200 requests; states built on the device)."""
    from wamp import serve

    return serve.poisson_workload(200, 1.0, 5, device_ic=True)


def test_served_stream_slice_equals_alone(gpu_ok, stream):
    """The first six requests of the stream plus its largest one, under Poisson
    arrivals (the slice arrives within ~2 s, while earlier requests are still
    resident); each served mixed result equals its solve alone."""
    from wamp import serve
    from wamp.solver import Solver

    big = max(stream, key=lambda r: r.cfg.grid.nx * r.cfg.grid.ny).rid
    reqs = stream[:6] + ([stream[big]] if big >= 6 else [])
    span = max(r.arrival for r in reqs)
    results, st = serve.serve_batched(serve.with_arrivals(reqs, span / 2.0))
    assert st["completed"] == len(reqs)
    by_id = {r.rid: r for r in results}
    big_req = next(r for r in reqs if r.rid == big)
    print(f"slice of {len(reqs)} requests, largest {big_req.cfg.name}: p50 "
          f"{st['p50_latency_s']:.2f} s, p99 {st['p99_latency_s']:.2f} s, "
          f"peak resident {st['peak_resident']}")
    assert st["peak_resident"] >= 2
    for r in reqs:
        s = Solver(r.cfg)
        s.set_state(r.state0())
        s.prepare()
        s.advance_to(r.t_end, log=None)
        alone, t = s.state(), s.time
        s.close()
        got = by_id[r.rid]
        assert got.t == t == pytest.approx(r.t_end, rel=1e-12)
        si, sj = r.cfg.grid.interior()
        assert np.array_equal(got.state[:, si, sj], alone[:, si, sj]), r.rid


def test_served_stream_small_requests_fp64_against_oracle(gpu_ok, stream):
    """The smallest droplet and triple-point requests of the stream, in FP64, to
    their end times, against the oracle's time loop."""
    from oracle.oracle import Oracle
    from wamp.config import PREC_FP64
    from wamp.solver import Solver

    picked = {}
    for r in sorted(stream, key=lambda r: (r.cfg.grid.nx * r.cfg.grid.ny, r.rid)):
        kind = r.cfg.name.split("_")[0]
        picked.setdefault(kind, r)
    assert len(picked) == 2, sorted(picked)
    for kind, r in picked.items():
        cfg = dataclasses.replace(r.cfg, precision=PREC_FP64)
        U0 = r.state0().cpu().numpy()
        orc = Oracle(cfg, threads=THREADS)
        orc.set_state(U0)
        orc.prepare()
        t_ref, n_ref = oracle_advance_to(orc, r.t_end)
        ref = orc.state()
        orc.close()
        s = Solver(cfg)
        s.set_state(U0)
        s.prepare()
        s.advance_to(r.t_end, log=None)
        got, t, n = s.state(), s.time, s.nstep
        s.close()
        errs = _interior_errs(cfg, got, ref)
        print(f"{cfg.name} to t={r.t_end} ({n} steps): rel L1", ["%.1e" % e for e in errs])
        assert n == n_ref and t == pytest.approx(t_ref, rel=1e-13)
        assert max(errs) <= 1e-9
