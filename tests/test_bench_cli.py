"""bench.py's command line on the CPU: multi-GPU requests are refused with a
message (this build runs on one GPU), and the options parse."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 2
    assert "one GPU only" in p.stderr
    assert p.stdout == ""


def test_help_lists_the_contract_options():
    p = subprocess.run([sys.executable, BENCH, "--help"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--cpu-steps", "--no-forms", "--form"):
        assert opt in p.stdout
