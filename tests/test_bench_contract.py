"""bench.py's command line: one GPU only, refused before any GPU work."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    for n in ("2", "8"):
        r = _bench("--gpus", n)
        assert r.returncode == 2
        assert "one GPU" in r.stderr and r.stdout == ""


def test_unknown_workload_is_a_usage_error():
    r = _bench("--workload", "c9")
    assert r.returncode == 2 and "--workload" in r.stderr
