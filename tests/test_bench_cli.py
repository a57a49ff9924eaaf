"""bench.py's command line on the CPU: more than one GPU is refused with a message (the bench
measures one MI355X), and every workload name is accepted by the parser."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    res = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 2
    assert "one GPU" in res.stderr
    assert res.stdout == ""


def test_workload_names():
    res = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0
    for w in ("c2", "c3", "c4", "c5", "c2fused", "c2flash", "c1periodic", "c2pm"):
        assert w in res.stdout
