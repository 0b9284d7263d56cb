"""bench.py's command-line contract that does not need a GPU."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def _bench():
    sys.path.insert(0, ROOT)
    import bench
    return bench


def test_roofline_of_a_one_launch_step_uses_the_timed_region():
    """A step that is one launch takes the kernel time from the timed region
    (step_ms); the mean of single-launch event pairs is kept beside it."""
    b = _bench()
    V = 256 ** 3
    name, roof = b.roofline({"fused_step_f32": 0.660}, V, "fused:p2u:64x4x128", step_ms=0.650)
    assert name == "fused_step_f32"
    assert roof["kernel_ms"] == 0.65 and roof["kernel_ms_single_launch_events"] == 0.66
    assert roof["kernel_ms_source"].startswith("timed-region")
    assert abs(roof["achieved"] - 184 * V / 0.650e-3 / 1e9) < 0.1
    assert abs(roof["frac"] - roof["achieved"] / roof["peak"]) < 1e-4
    # traffic of that schedule from the committed PMC profile
    assert roof["traffic"] is not None and roof["traffic"] > 184 * V


def test_roofline_of_a_multi_launch_step_uses_the_launch_events():
    b = _bench()
    V = 256 ** 3
    name, roof = b.roofline({"qstep_f64": 1.05, "force_lbm_f64": 1.82}, V, "two-kernel", step_ms=2.87)
    assert name == "force_lbm_f64"
    assert roof["kernel_ms"] == 1.82 and roof["kernel_ms_source"].startswith("event pairs")
    assert roof["traffic"] is None


def test_fp64_roofline_finds_the_traffic_of_its_schedule():
    b = _bench()
    V = 256 ** 3
    _, roof = b.roofline({"fused_step_f64": 2.15}, V, "fused:fp64:16x8x64", step_ms=2.14)
    assert roof["traffic"] is not None and roof["traffic"] > 368 * V
    _, roof = b.roofline({"fused_step_f64": 2.15}, V, "fused:fp64:no-such-tile", step_ms=2.14)
    assert roof["traffic"] is None
