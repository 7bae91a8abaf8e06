"""bench.py without a GPU: more than one GPU is refused with a message, and with no device the bench fails loudly
(no CPU fallback measures anything)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(*args, timeout=300):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT, capture_output=True,
                          text=True, timeout=timeout)


@pytest.mark.parametrize("n", ["2", "8"])
def test_more_than_one_gpu_is_refused(n):
    r = run_bench("--gpus", n)
    assert r.returncode != 0
    assert "ONE GPU" in (r.stdout + r.stderr) and f"--gpus {n}" in (r.stdout + r.stderr)
    assert not any(line.startswith("{") for line in r.stdout.splitlines())  # no result line


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present: the bench would run")
def test_no_device_fails_loudly():
    r = run_bench("--steps", "1", "--warmup", "0", "--no-cpu")
    assert r.returncode != 0
    assert "MLEM_ERR_NO_DEVICE" in r.stderr
    assert not any(line.startswith("{") for line in r.stdout.splitlines())
