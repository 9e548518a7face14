"""bench.py on the CPU: a run on more than one GPU is refused with a message, before any device
work (the build runs on one GPU)."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2, (r.returncode, r.stdout, r.stderr)
    line = [x for x in r.stdout.splitlines() if x.startswith("{")][-1]
    msg = json.loads(line)["error"]
    assert "one GPU" in msg
