"""bench.py's command-line contract on the host: more than one GPU is refused with a message and a
non-zero status before any GPU work, and an unknown workload is refused by name."""

import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def run(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_unknown_workload_is_refused():
    r = run("--workload", "no_such_workload")
    assert r.returncode != 0
    assert "unknown workload no_such_workload" in (r.stderr + r.stdout)
