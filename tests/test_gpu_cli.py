"""End to end through the command-line front end (the reference's argv
contract) and through a minimal ctypes binding of the C ABI, on the 2D ring
workload, against the FP64 oracle."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import mlem_amd as M
import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "bachelorarbeit-mlem_amd", "bin", "mlem")


@pytest.fixture(scope="module")
def files(c1, tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    A, g = c1["A"], c1["g"]
    mp, gp = str(d / "ring.csr4"), str(d / "ring.i32")
    M.write_csr4(mp, A)
    M.write_raw(gp, g)
    s32 = M.colsum_f32(A)
    f0, _ = M.initial_image(g, A.rows, s32)
    return dict(dir=d, matrix=mp, sino=gp, s32=s32, f0=f0)


def run(*args, timeout=300):
    print("cli:", " ".join(args), flush=True)
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("trans,spmv,extra", [(0, 0, []), (0, 1, []), (1, 0, []), (1, 1, []),
                                              (0, 0, ["--transpose=gpu", "--graph=1"])])
def test_cli_reconstruction_matches_oracle(c1, files, trans, spmv, extra):
    A, g = c1["A"], c1["g"]
    out = str(files["dir"] / f"img_{trans}{spmv}{len(extra)}.f32")
    r = run(files["matrix"], files["sino"], "20", "1", "6", "7", str(trans), str(spmv), f"--out={out}",
            "--sens=host32", *extra)
    assert r.returncode == 0, r.stderr
    img = M.read_raw(out, np.float32)
    ref = O.mlem_f64(A.rowptr, A.col, A.val, A.rows, A.cols, g, files["s32"].astype(np.float64), files["f0"], 20)
    err = np.linalg.norm(img - ref) / np.linalg.norm(ref)
    assert err <= 1e-5, err
    s = np.float32(0)
    for v in img:
        s = np.float32(s + v)
    assert r.stdout.endswith(f"\nSum f: {s:f}\n")


def test_cli_16bit_values(c1, files):
    """--values=uf16: the blocked kernels on the quantized matrix (its own FP64 sensitivity), within 1e-4 of the
    FP64 oracle on the f32 matrix after 20 iterations; a bad value name prints the help."""
    A, g = c1["A"], c1["g"]
    out = str(files["dir"] / "img_uf16.f32")
    r = run(files["matrix"], files["sino"], "20", "1", "6", "7", "0", "0", f"--out={out}", "--values=uf16")
    assert r.returncode == 0, r.stderr
    img = M.read_raw(out, np.float32)
    s64 = O.colsum_f64(A.rowptr, A.col, A.val, A.rows, A.cols)
    f0, _ = M.initial_image(g, A.rows, s64.astype(np.float32))
    ref = O.mlem_f64(A.rowptr, A.col, A.val, A.rows, A.cols, g, s64.astype(np.float32).astype(np.float64), f0, 20)
    assert np.linalg.norm(img - ref) / np.linalg.norm(ref) <= 1e-4
    r = run(files["matrix"], files["sino"], "2", "1", "2", "2", "0", "0", "--values=int8")
    assert r.returncode == 1 and "--values" in r.stdout


@pytest.mark.parametrize("values", ["f32", "uf16"])
def test_cli_resident_tiles_equals_resident_csr(c1, files, values):
    """--resident=tiles (no CSR copies on the GPU): the image is bitwise the one of the resident-CSR blocked path
    (the same tiles; x0 = 1, since a tiles-only matrix with 16-bit values reports the quantized matrix's
    sensitivity, from which the reference's initial image is scaled); argv[7] = 1 or argv[8] = 1 is refused."""
    outs = []
    for res in ("csr", "tiles"):
        out = str(files["dir"] / f"img_res_{res}_{values}.f32")
        r = run(files["matrix"], files["sino"], "15", "1", "8", "8", "0", "0", f"--out={out}", "--transpose=gpu",
                f"--resident={res}", f"--values={values}", "--x0=one")
        assert r.returncode == 0, r.stderr
        outs.append(M.read_raw(out, np.float32))
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
    r = run(files["matrix"], files["sino"], "2", "1", "2", "2", "0", "1", "--resident=tiles")
    assert r.returncode == 1


def test_cli_streamed_load_equals_host_path(c1, files):
    # --transpose=gpu with the GPU sensitivity streams the csr4 file to the
    # device (no host CSR); the image is bitwise the one of the host path.
    outs = []
    for mode in ("host", "gpu"):
        out = str(files["dir"] / f"img_{mode}.f32")
        r = run(files["matrix"], files["sino"], "15", "1", "8", "8", "0", "0", f"--out={out}", f"--transpose={mode}",
                "--verbose")
        assert r.returncode == 0, r.stderr
        outs.append(M.read_raw(out, np.float32))
        if mode == "gpu":
            assert "streamed" in r.stderr
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


def test_cli_file_errors(c1, files):
    r = run("/nonexistent.csr4", files["sino"], "2", "1", "2", "2", "0", "0")
    assert r.returncode == 2
    short = str(files["dir"] / "short.i32")
    M.write_raw(short, c1["g"][:10])
    r = run(files["matrix"], short, "2", "1", "2", "2", "0", "0")
    assert r.returncode == 3


def test_integration_ctypes_stub(c1):
    """A minimal ctypes binding of the C ABI (what a maintainer would add to call mlem_run)."""
    hip = C.CDLL(os.path.join(ROOT, "bachelorarbeit-mlem_amd", "lib", "libmlem_hip.so"))
    hip.mlem_last_error.restype = C.c_char_p

    class Opts(C.Structure):
        _fields_ = [(n, C.c_uint32) for n in ("spmv", "using_trans", "secsize_fp", "secsize_bp", "fp_lanes",
                                              "bp_lanes", "profile", "use_graph", "bp_spmv", "tile_width",
                                              "tile_rows", "tile_groups", "values")]

    assert C.sizeof(Opts) == C.sizeof(M.Opts)
    P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731

    def reconstruct(rowptr, col, val, rows, cols, g, iters):
        m = C.c_void_p()
        assert hip.mlem_matrix_create(0, rows, cols, C.c_uint64(len(col)), P(rowptr), P(col), P(val),
                                      None, None, None, C.byref(m)) == 0, hip.mlem_last_error()
        assert hip.mlem_matrix_build_transpose(m) == 0
        o = Opts()
        hip.mlem_opts_default(C.byref(o))
        f0 = np.ones(cols, np.float32)
        out = np.empty(cols, np.float32)
        assert hip.mlem_run(m, P(g), None, P(f0), P(out), iters, C.byref(o)) == 0, hip.mlem_last_error()
        hip.mlem_matrix_destroy(m)
        return out

    A, g = c1["A"], c1["g"]
    img = reconstruct(A.rowptr, A.col, A.val, A.rows, A.cols, g, 10)
    s64 = O.colsum_f64(A.rowptr, A.col, A.val, A.rows, A.cols)
    ref = O.mlem_f64(A.rowptr, A.col, A.val, A.rows, A.cols, g, s64.astype(np.float32).astype(np.float64),
                     np.ones(A.cols), 10)
    assert np.linalg.norm(img - ref) / np.linalg.norm(ref) <= 1e-5
