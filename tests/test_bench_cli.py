"""bench.py's command line on the CPU: the single-GPU contract and the flags
that the profiling commands use (no GPU is touched)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_exits_with_message():
    r = run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_help_lists_flags():
    r = run("--help")
    assert r.returncode == 0
    for flag in ("--steps", "--warmup", "--split", "--offsets", "--no-large", "--no-alt", "--no-cpu", "--no-plain",
                 "--exact-iters"):
        assert flag in r.stdout, flag


def test_offsets_flag_parses_a_list():
    """--offsets takes comma-separated milliseconds (the start offsets tried
    for the concurrent sub-batches); the single-GPU check still comes first."""
    r = run("--offsets", "0,1.5,3", "--gpus", "2")
    assert r.returncode == 2 and "one GPU" in r.stderr
    r = run("--offsets", "zero")
    assert r.returncode == 2 and "--offsets" in r.stderr
