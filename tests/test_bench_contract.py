"""bench.py pieces that run without a GPU: the one-GPU guard and the lookup of
the committed rocprof traffic figure for the roofline."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr


def test_traffic_record_lookup(tmp_path, monkeypatch):
    sys.path.insert(0, ROOT)
    import bench

    v = bench.measured_traffic("mach10_cylinder_4096x2048", "k_faces")
    assert isinstance(v, int) and v > 1_074_000_000  # above the algorithmic 128 B/cell
    assert bench.measured_traffic("mach10_cylinder_512x256", "k_faces") is None
    # a record for another workload is not used
    prof = tmp_path / "profiles"
    prof.mkdir()
    (prof / "k_faces_traffic_x_1x1.txt").write_text("workload=other\nkernel=k_faces\ntraffic_bytes=5\n")
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    assert bench.measured_traffic("x_1x1", "k_faces") is None


def test_limiter_rates_from_the_committed_record():
    sys.path.insert(0, ROOT)
    import bench

    # the FP64 limiter at this size reads 14 face planes (112 B), U^k and U^n
    # (96 B) and writes U^(k+1) (48 B): 256 B per cell
    r = bench.limiter_rates("mach10_cylinder_4096x2048", "k_lu_lo", 0.861, 6000.0)
    assert r["kernel"] == "k_lu_lo"
    assert r["alg_bytes"] == 256 * 4096 * 2048
    assert r["traffic"] > r["alg_bytes"]
    assert abs(r["traffic_gbs"] - r["traffic"] / 0.861e-3 / 1e9) < 0.1
    assert abs(r["traffic_frac_of_copy"] - r["traffic_gbs"] / 6000.0) < 1e-4
    assert bench.limiter_rates("mach10_cylinder_512x256", "k_lu_lo", 0.1, 6000.0)["traffic"] is None
    assert bench.limiter_rates("mach10_cylinder_4096x2048", "k_lu_march", 0.8, 6000.0)["traffic"] is None


def test_cpu_device_leg_runs_on_the_host():
    sys.path.insert(0, ROOT)
    import bench
    from wamp import cases

    cfg, U0 = cases.mach10_cylinder(64, 32)
    r = bench.cpu_device_leg(cfg, U0, 2)
    assert r["value"] > 0 and r["unit"] == "Gcell-updates/s" and r["threads"] >= 1
    assert "64x32" in r["sample"]
