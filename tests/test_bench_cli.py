"""bench.py's command line on the CPU: it measures one GPU and refuses
--gpus N > 1 with a message (no GPU is touched), and --help lists the
workloads."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_refuses_more_than_one_gpu():
    p = _run("--gpus", "2")
    assert p.returncode == 2
    assert "one GPU" in (p.stdout + p.stderr)


def test_help_lists_workloads():
    p = _run("--help")
    assert p.returncode == 0
    for w in ("heightfield_1m", "trace", "build", "batch64", "rows", "light_tree", "subsurface"):
        assert w in p.stdout, w
