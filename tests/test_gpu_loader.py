"""The streamed csr4 loader (mlem_matrix_load_csr4): bit-exact device arrays,
chunk and buffer-rotation boundaries, bad columns, and reconstructions from a
loaded matrix equal to those from an uploaded one."""
import os
import struct

import numpy as np
import pytest

import mlem_amd as M
from matrices import from_lengths, make

pytestmark = pytest.mark.gpu


def same(a, b):
    return a.shape == b.shape and np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))


@pytest.mark.parametrize("kind", ["ragged", "empty_rows", "one_huge_row", "all_empty", "single_row", "tiny"])
def test_load_matches_source(tmp_path, kind):
    A = make(kind, seed=41)
    p = str(tmp_path / f"{kind}.csr4")
    M.write_csr4(p, A)
    print(f"load {kind}", flush=True)
    m, st = M.ResidentMatrix.load_csr4(p, threads=3)
    with m:
        B = m.download()
        assert same(B.rowptr, A.rowptr) and same(B.col, A.col) and same(B.val, A.val)
        assert st["bytes"] == 8 * A.nnz


def test_load_crosses_chunks_and_buffers(tmp_path):
    # ~85M entries = 680 MB: six 128 MiB chunks, the four staging buffers reused
    rng = np.random.default_rng(5)
    lens = rng.integers(0, 2000, 85000)
    A = from_lengths(lens, 1 << 20, seed=6, dup=True)
    p = str(tmp_path / "big.csr4")
    M.write_csr4(p, A)
    for threads in (1, 7):
        print(f"load big threads={threads}", flush=True)
        m, st = M.ResidentMatrix.load_csr4(p, threads=threads)
        with m:
            B = m.download()
            assert same(B.rowptr, A.rowptr) and same(B.col, A.col) and same(B.val, A.val)
        print(f"  {st['bytes'] / 1e6:.0f} MB in {st['stream_s'] * 1e3:.1f} ms", flush=True)


def test_load_rejects_bad_column(tmp_path):
    A = make("ragged", seed=42)
    p = str(tmp_path / "bad.csr4")
    M.write_csr4(p, A)
    raw = bytearray(open(p, "rb").read())
    scb = struct.unpack_from("<I", raw, 24)[0]
    pos = 28 + scb + 8 * A.rows + 8 * (A.nnz // 2)
    struct.pack_into("<I", raw, pos, A.cols + 3)
    open(p, "wb").write(bytes(raw))
    with pytest.raises(M.MlemError):
        M.ResidentMatrix.load_csr4(p)


def test_reconstruction_from_loaded_matrix(tmp_path, c1):
    A, g = c1["A"], c1["g"]
    p = str(tmp_path / "ring.csr4")
    M.write_csr4(p, A)
    m, _ = M.ResidentMatrix.load_csr4(p)
    f0 = np.ones(A.cols, np.float32)
    with m:
        m.build_transpose()
        s32, _ = m.sensitivity()
        a = M.mlem_run(m, g, f0, 10, sens=s32)
    with M.ResidentMatrix(A) as u:
        u.build_transpose()
        b = M.mlem_run(u, g, f0, 10, sens=s32)
    assert same(a, b)
