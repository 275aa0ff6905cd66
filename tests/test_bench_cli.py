"""bench.py's command-line contract that needs no GPU: more than one GPU is
refused with a message (one GPU's share is what this build measures)."""
import json
import os
import subprocess
import sys

from conftest import ROOT


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]
