"""bench.py's command line on the CPU: more than one GPU is refused with a
message and exit status 2 (there is no multi-GPU path), before any device
work; --help lists the options the measurements use."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


def test_help_lists_the_measurement_options():
    r = run("--help")
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--config", "--stream-ceiling", "--copy-calibration", "--serve-window",
                "--cpu-warmup", "--no-c4"):
        assert opt in r.stdout, opt
