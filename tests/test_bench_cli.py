"""bench.py's command line on a machine without a GPU: more than one GPU is
refused with a message before anything touches a device, and --help lists
the modes."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def run(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = run("--gpus", "2")
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_help_lists_the_modes():
    r = run("--help")
    assert r.returncode == 0
    for mode in ("matched", "plain", "mixed", "ira"):
        assert mode in r.stdout
