"""bench.py's command-line contract, checked without a GPU: it measures one
GPU only and says so for any larger count; its options parse."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args, env=None):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, env=env)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = _bench("--gpus", "8", "--steps", "1", "--warmup", "0")
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


def test_world_size_above_one_is_refused():
    env = dict(os.environ, WORLD_SIZE="2")
    r = _bench("--steps", "1", env=env)
    assert r.returncode == 2 and "one GPU only" in r.stderr


def test_help_lists_the_modes():
    r = _bench("--help")
    assert r.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--multi-only", "--serve", "--index-stress", "--layout"):
        assert opt in r.stdout
