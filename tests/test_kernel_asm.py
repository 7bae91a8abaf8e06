"""The gfx950 assembly of the LDS-tiled kernels (tiled.hip), compiled here:
the segmented carry scans issue their DPP steps from inline asm, which the
compiler's hazard recognizer does not look into.  Each such step carries its
own s_nop 1 (two wait states after a VALU write of the register it reads);
the other DPP hazard, a VALU write of EXEC (v_cmpx) within five wait states,
cannot arise only as long as the file contains no v_cmpx.  Both are checked
on the compiler's output."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "bachelorarbeit-mlem_amd")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def tiled_asm(tmp_path_factory):
    if not (os.path.exists(HIPCC) or shutil.which("hipcc")):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("asm") / "tiled.s")
    cmd = [HIPCC if os.path.exists(HIPCC) else "hipcc", "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
           f"-I{ROOT}/include", f"-I{PKG}/csrc", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-munsafe-fp-atomics", "--cuda-device-only", "-S", f"{PKG}/csrc/kernels/tiled.hip", "-o", out]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    with open(out) as f:
        return f.read().splitlines()


def test_no_valu_writes_of_exec(tiled_asm):
    assert not [l for l in tiled_asm if re.match(r"\s*v_cmpx", l)]


def test_inline_dpp_steps_wait_after_valu_writes(tiled_asm):
    steps = 0
    for i, line in enumerate(tiled_asm):
        if re.match(r"\s*v_(fmac|mul)_f32_dpp\b", line):
            steps += 1
            prev = [l for l in tiled_asm[max(0, i - 3):i] if l.strip() and not l.strip().startswith(";")]
            assert prev and re.match(r"\s*s_nop\s+[1-9]", prev[-1]), (i, line)
    assert steps > 0
