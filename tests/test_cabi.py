"""The C ABI (include/*.h) without a GPU: both libraries load, export every
declared function and nothing else under the mlem_ prefix; the product path
does not touch the oracle; the command-line front end checks its arguments
like main.cu:42-62 before touching a GPU."""
import ctypes as C
import os
import re
import subprocess

import pytest

import mlem_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "bachelorarbeit-mlem_amd")
LIBS = {"mlem_host.h": os.path.join(PKG, "lib", "libmlem_host.so"),
        "mlem_hip.h": os.path.join(PKG, "lib", "libmlem_hip.so")}
CLI = os.path.join(PKG, "bin", "mlem")


def declared(header):
    s = open(os.path.join(ROOT, "include", header)).read()
    s = re.sub(r"/\*.*?\*/", "", s, flags=re.S)
    s = re.sub(r"//[^\n]*", "", s)
    return set(re.findall(r"(?:^|[;}\s])(?:const\s+)?(?:unsigned\s+)?\w+[\s\*]+\**(mlem_\w+)\s*\(", s, flags=re.M))


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], check=True, capture_output=True, text=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.split() and ln.split()[-1].startswith("mlem_")}


@pytest.mark.parametrize("header", sorted(LIBS))
def test_library_exports_every_declared_function(header):
    lib = LIBS[header]
    assert os.path.exists(lib), f"{lib} is not built (run __graft_entry__.build())"
    names = declared(header)
    assert len(names) >= 20
    dll = C.CDLL(lib)
    for n in sorted(names):
        assert hasattr(dll, n), f"{n} declared in include/{header} but not exported by {os.path.basename(lib)}"
    assert exported(lib) == names, "exported mlem_* symbols differ from the header's declarations"


def test_status_strings_and_version_without_gpu():
    hip = M.hip_lib()
    assert hip.mlem_status_string(0) and hip.mlem_status_string(4)
    assert M.version()
    assert M.host_lib().mlem_host_status_string(7)


def test_product_path_does_not_use_the_oracle():
    for lib in LIBS.values():
        needed = subprocess.run(["objdump", "-p", lib], check=True, capture_output=True, text=True).stdout
        assert "oracle" not in needed
    for dirpath, _, files in os.walk(os.path.join(PKG, "mlem_amd")):
        for f in files:
            if f.endswith(".py"):
                src = open(os.path.join(dirpath, f)).read()
                assert not re.search(r"^\s*(import|from)\s+oracle", src, flags=re.M), f


def run_cli(*args):
    assert os.path.exists(CLI), f"{CLI} is not built (run __graft_entry__.build())"
    env = dict(os.environ, HIP_VISIBLE_DEVICES=os.environ.get("HIP_VISIBLE_DEVICES", ""))
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120, env=env)


@pytest.mark.parametrize("args", [
    [],
    ["m.csr4", "g.i32", "10", "1", "2", "2", "0"],             # 7 positional
    ["m.csr4", "g.i32", "0", "1", "2", "2", "0", "0"],         # iterations 0
    ["m.csr4", "g.i32", "10", "0", "2", "2", "0", "0"],        # nGPU 0
    ["m.csr4", "g.i32", "10", "1", "0", "2", "0", "0"],        # secsize_fp 0 (checked even if unused)
    ["m.csr4", "g.i32", "10", "1", "2", "0", "0", "0"],        # secsize_bp 0
    ["m.csr4", "g.i32", "10", "1", "2", "2", "2", "0"],        # trans not 0/1
    ["m.csr4", "g.i32", "10", "1", "2", "2", "0", "3"],        # spmv not 0/1
    ["m.csr4", "g.i32", "x", "1", "2", "2", "0", "0"],         # strtol -> 0
    ["m.csr4", "g.i32", "10", "1", "2", "2", "0", "0", "extra"],
    ["m.csr4", "g.i32", "10", "1", "2", "2", "0", "0", "--bogus=1"],
    ["m.csr4", "g.i32", "10", "1", "2", "2", "0", "0", "--sens=fp16"],
])
def test_cli_usage_errors_print_help(args):
    r = run_cli(*args)
    assert r.returncode == 1
    assert "Wrong parameters!" in r.stdout and "argv[8]" in r.stdout


def test_cli_more_gpus_than_present_prints_help():
    r = run_cli("m.csr4", "g.i32", "10", "4096", "2", "2", "0", "0")
    assert r.returncode == 1 and "Wrong parameters!" in r.stdout
