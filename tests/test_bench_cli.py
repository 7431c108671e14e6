"""bench.py's command line, without a GPU: the one-GPU guard and the flags
that size its sections."""
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(*args):
    return subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=REPO)


def test_more_than_one_gpu_is_refused():
    # this build measures one GPU: N > 1 ends with a message, before any GPU work
    p = run_bench("--gpus", "2")
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""


def test_section_flags_listed():
    p = run_bench("--help")
    assert p.returncode == 0
    for flag in ("--steps", "--warmup", "--no-cpu", "--no-batched", "--no-served", "--no-stress", "--lam", "--slots",
                 "--stress-solves", "--stress-iters"):
        assert flag in p.stdout
