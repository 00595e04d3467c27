"""The benchmark's command line on the CPU: a request for more than one GPU
ends with status 2 and one JSON line naming the reason, before anything
touches a device."""
import json
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_json_line():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], cwd=REPO,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2, r.stderr
    lines = [l for l in r.stdout.splitlines() if l.strip()]
    assert len(lines) == 1
    rec = json.loads(lines[0])
    assert rec["n_gpus"] == 2
    assert "one GPU" in rec["error"]
