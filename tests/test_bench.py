"""bench.py's contract on the CPU: it measures one GPU and says so for any
other count, and its CPU baseline (the oracle on a bounded sample of the
same workload) reports the fields the JSON line carries."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_cpu_baseline_reports_its_sample():
    """Both CPU numbers (the torch restatement on every host thread, and the
    NumPy oracle on one core) on the same strip of the throughput workload,
    finite (the baseline asserts it) and with the sample's configuration."""
    sys.path.insert(0, ROOT)
    import bench
    b = bench.cpu_baseline(rows=2, np_steps=1, torch_steps=1)
    assert b["value"] > 0 and b["unit"] == "particle-steps/s"
    assert b["cores"] == len(os.sched_getaffinity(0)) and b["kind"] == "port"
    assert "1024x2-cell strip" in b["sample"] and "side rollers" in b["sample"]
    assert b["config"]["particles"] == 4 * 1024 * 2
    s = b["single_core"]
    assert s["value"] > 0 and s["cores"] == 1 and s["kind"] == "port"
