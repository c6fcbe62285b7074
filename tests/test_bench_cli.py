"""bench.py's command line on a machine without a GPU: more than one GPU is
refused with a JSON message and status 2 (the path is measured on one GPU),
and --help lists the section switches."""
import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       cwd=ROOT, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_help_lists_the_sections():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--help"], capture_output=True, text=True, cwd=ROOT,
                       timeout=120)
    assert r.returncode == 0
    for flag in ("--steps", "--warmup", "--no-traversal", "--no-batched", "--no-serve", "--no-autotune"):
        assert flag in r.stdout
