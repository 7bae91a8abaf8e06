"""bench.py without a GPU: more than one GPU is refused with a message, with no device the bench fails loudly
(no CPU fallback measures anything), and the defaults that depend on the workload resolve as documented."""
import json
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


def resolved(*argv):
    """bench.parse_args(argv) in a child process (importing bench sets the OpenMP binding of its process)."""
    code = f"import bench, json; print(json.dumps(vars(bench.parse_args({list(argv)!r}))))"
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


@pytest.mark.parametrize("argv,want", [
    ((), dict(workload="c2", values="f32", graph=0, steps=100, autotune=0, fp="merge", bp="merge")),
    (("--workload", "c1"), dict(values="f32", graph=1, steps=2000, autotune=0, fp="merge", bp="vector")),
    (("--workload", "c1", "--graph", "0", "--steps", "50", "--bp", "merge"), dict(graph=0, steps=50, bp="merge")),
    (("--batch", "32"), dict(values="uf16", graph=1, steps=100, autotune=1, bp="merge")),
    (("--workload", "c1", "--batch", "8"), dict(graph=1, steps=100, autotune=1, bp="merge")),
    (("--serve", "c1,mid,c2"), dict(values="uf16", graph=1, autotune=0, bp="merge")),
])
def test_workload_defaults(argv, want):
    got = resolved(*argv)
    assert {k: got[k] for k in want} == want
