"""bench.py's command line without a GPU: the one-GPU contract and the flags of its profiled runs."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    r = run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_help_lists_the_profiling_flags():
    r = run("--help")
    assert r.returncode == 0
    for flag in ("--steps", "--warmup", "--path", "--no-overlap", "--events-only", "--no-cpu-baseline", "--c3",
                 "--copy-probe"):
        assert flag in r.stdout, flag
