"""bench.py's command-line contract that needs no GPU: it measures one GPU and
refuses more with a message and exit status 2, before touching the device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    for n in (2, 8):
        r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)], capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 2 and "measures one GPU" in r.stderr and r.stdout == "", (n, r.stdout, r.stderr)
