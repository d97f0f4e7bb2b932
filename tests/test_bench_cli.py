"""bench.py's command line on a machine without a GPU: more than one GPU is
refused with exit status 2 and a message, before any device is touched."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    for n in (2, 8):
        r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)],
                           capture_output=True, text=True, timeout=120, cwd=ROOT)
        assert r.returncode == 2, (r.returncode, r.stderr)
        assert "one GPU only" in r.stderr and r.stdout == ""
