"""bench.py's command-line contract, on the CPU: one GPU only (a request
for more exits with a message before anything touches a device), and every
mode the help text lists."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args, env=None):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT, capture_output=True,
                          text=True, timeout=300, env=env)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_world_size_above_one_is_refused():
    env = dict(os.environ, WORLD_SIZE="2")
    r = _bench(env=env)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_help_lists_every_mode():
    r = _bench("--help")
    assert r.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup", "--core", "--kernels", "--kernels2d", "--band", "--calibrate-copy",
                 "--no-cpu", "--no-compare"):
        assert flag in r.stdout, flag
