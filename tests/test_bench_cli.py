"""bench.py's command line on the host: it measures one GPU and says so when
asked for more (no GPU is touched before that check)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    res = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 2
    assert "--gpus 2 is not supported" in res.stderr
    assert res.stdout == ""  # no JSON line
