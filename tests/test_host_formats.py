"""csr4 matrix files, raw vectors and the AoS -> SoA CSR build (host library,
no GPU).  The tiny csr4 fixture was assembled byte by byte from the format
(tests/golden/make_golden.py); the error cases follow the checks of
Csr4Matrix::Csr4Matrix (csr4matrix.cpp:44-107,141-148)."""
import json
import os
import struct

import numpy as np
import pytest

import mlem_amd as M
import oracle as O
from matrices import KINDS, make

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_bytes(name):
    with open(os.path.join(GOLDEN, name), "rb") as f:
        return f.read()


def test_tiny_csr4_fixture_decodes():
    exp = json.load(open(os.path.join(GOLDEN, "tiny_3x4.json")))
    with M.Csr4File(os.path.join(GOLDEN, "tiny_3x4.csr4")) as f:
        assert (f.rows, f.cols, f.nnz, f.flags, f.scan_config_bytes) == (
            exp["rows"], exp["cols"], exp["nnz"], exp["flags"], exp["scan_config_bytes"])
        assert list(f.row_ends()) == exp["row_ends"]
        assert [f.elements_in_row(r) for r in range(f.rows)] == list(np.diff([0] + exp["row_ends"]))
        sc = f.scanner()
        e = exp["scanner"]
        assert sc.n_layers == e["n_layers"] and sc.n_blocks == e["n_blocks"] and sc.blocksize == e["blocksize"]
        assert sc.n_rings == e["n_rings"] and list(sc.grid) == e["grid"]
        np.testing.assert_array_equal(np.array(sc.layer[0], np.float32), np.array(e["layer0"], np.float32))
        np.testing.assert_array_equal(np.array(sc.voxel_size, np.float32), np.array(e["voxel_size"], np.float32))
        for threads in (1, 2, 3):
            A = f.to_csr(threads)
            assert list(A.rowptr) == exp["rowptr"] and list(A.col) == exp["col"]
            np.testing.assert_array_equal(A.val, np.array(exp["val"], np.float32))
        np.testing.assert_array_equal(A.to_scipy().toarray(), np.array(exp["dense"], np.float32))
    # the oracle's restatement of csr_format_for_cuda over the same bytes
    raw = read_bytes("tiny_3x4.csr4")
    rows = exp["rows"]
    ends = np.frombuffer(raw, np.uint64, rows, 28 + 60)
    rp, col, val = O.csr_from_csr4(rows, ends, raw[28 + 60 + 8 * rows:])
    assert list(rp) == exp["rowptr"] and list(col) == exp["col"]


def test_writer_reproduces_the_hand_assembled_file(tmp_path):
    exp = json.load(open(os.path.join(GOLDEN, "tiny_3x4.json")))
    A = M.Csr(3, 4, np.array(exp["rowptr"], np.uint64), np.array(exp["col"], np.uint32),
              np.array(exp["val"], np.float32))
    with M.Csr4File(os.path.join(GOLDEN, "tiny_3x4.csr4")) as f:
        sc = f.scanner()
    p = str(tmp_path / "w.csr4")
    M.write_csr4(p, A, scanner=sc)
    assert open(p, "rb").read() == read_bytes("tiny_3x4.csr4")


def test_csr4_round_trip_and_csr_build(tmp_path):
    for k in KINDS:
        A = make(k, seed=31)
        p = str(tmp_path / f"{k}.csr4")
        M.write_csr4(p, A, flags=5)
        with M.Csr4File(p) as f:
            assert (f.rows, f.cols, f.nnz, f.flags) == (A.rows, A.cols, A.nnz, 5)
            ends = f.row_ends()
            assert np.array_equal(ends, A.rowptr[1:])
            for r in range(0, A.rows, max(1, A.rows // 17)):
                assert f.elements_in_row(r) == int(A.rowptr[r + 1] - A.rowptr[r])
            for t in (1, 4, 8):
                B = f.to_csr(t)
                assert np.array_equal(B.rowptr, A.rowptr) and np.array_equal(B.col, A.col)
                assert np.array_equal(B.val.view(np.uint32), A.val.view(np.uint32))
        # brute force of the AoS -> SoA build from the raw bytes
        raw = open(p, "rb").read()
        scb = struct.unpack_from("<I", raw, 24)[0]
        data = np.frombuffer(raw, np.dtype([("c", "<u4"), ("v", "<f4")]), A.nnz, 28 + scb + 8 * A.rows)
        assert np.array_equal(data["c"], A.col) and np.array_equal(data["v"].view(np.uint32), A.val.view(np.uint32))


def test_csr4_scanner_block_round_trip(tmp_path):
    sc = M.Scanner()
    sc.n_layers = 2
    for L in range(2):
        for q in range(5):
            sc.layer[L][q] = 1.5 * L + q
    sc.n_blocks, sc.blocksize, sc.n_rings = 36, 4, 3
    for q in range(3):
        sc.block_gap[q] = 0.25 * q
    for q in range(2):
        sc.ring_gap[q] = 2.0 + q
    sc.voxel_size[:] = [0.5, 0.5, 0.75]
    sc.grid[:] = [10, 20, 30]
    A = make("tiny")
    p = str(tmp_path / "s.csr4")
    M.write_csr4(p, A, scanner=sc)
    with M.Csr4File(p) as f:
        assert f.scan_config_bytes == 4 + 2 * 20 + 4 + 4 + 3 * 4 + 4 + 2 * 4 + 12 + 12
        s2 = f.scanner()
    assert s2.n_layers == 2 and s2.n_blocks == 36 and s2.blocksize == 4 and s2.n_rings == 3
    assert [s2.layer[1][q] for q in range(5)] == [1.5 + q for q in range(5)]
    assert [s2.block_gap[q] for q in range(3)] == [0.0, 0.25, 0.5] and [s2.ring_gap[q] for q in range(2)] == [2.0, 3.0]
    assert list(s2.grid) == [10, 20, 30] and list(s2.voxel_size) == [0.5, 0.5, 0.75]


def corrupt(tmp_path, name, mutate):
    raw = bytearray(read_bytes("tiny_3x4.csr4"))
    raw = mutate(raw)
    p = str(tmp_path / name)
    with open(p, "wb") as f:
        f.write(bytes(raw))
    return p


@pytest.mark.parametrize("case", ["magic", "version", "short_header", "truncated", "scan_short",
                                  "row_ends_decrease", "row_end_past_nnz", "layers_overflow"])
def test_csr4_rejects_malformed_files(tmp_path, case):
    def m(raw):
        if case == "magic":
            raw[0:4] = b"XCSR"
        elif case == "version":
            raw[4] = 3
        elif case == "short_header":
            raw = raw[:20]
        elif case == "truncated":
            raw = raw[:-3]
        elif case == "scan_short":
            struct.pack_into("<I", raw, 24, 56)
        elif case == "row_ends_decrease":
            struct.pack_into("<Q", raw, 28 + 60 + 8, 1)  # row 1 ends before row 0
        elif case == "row_end_past_nnz":
            struct.pack_into("<Q", raw, 28 + 60 + 16, 9)
        elif case == "layers_overflow":
            struct.pack_into("<I", raw, 28, 1000)
        return raw
    p = corrupt(tmp_path, case, m)
    with pytest.raises(M.MlemError):
        M.Csr4File(p)


def test_csr4_accepts_trailing_bytes(tmp_path):
    # The size check asks for at least the bytes the header implies
    # (csr4matrix.cpp:95-102), so trailing data is accepted.
    p = corrupt(tmp_path, "long", lambda raw: raw + b"\0" * 8)
    with M.Csr4File(p) as f:
        assert f.nnz == 5 and list(f.to_csr().col) == [1, 3, 0, 2, 3]


def test_csr4_missing_file():
    with pytest.raises(M.MlemError):
        M.Csr4File("/nonexistent/dir/none.csr4")


def test_raw_vectors(tmp_path):
    g = M.read_raw(os.path.join(GOLDEN, "sino_5.i32"), np.int32)
    assert list(g) == [3, 0, -7, 2147483647, 12]
    rng = np.random.default_rng(0)
    for dt in (np.int32, np.float32):
        a = (rng.random(1001) * 1e6).astype(dt)
        p = str(tmp_path / f"v.{np.dtype(dt).name}")
        M.write_raw(p, a)
        assert os.path.getsize(p) == 4 * a.size
        b = M.read_raw(p, dt)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    p = str(tmp_path / "odd")
    with open(p, "wb") as f:
        f.write(b"\1\2\3\4\5\6")
    with pytest.raises(M.MlemError):
        M.read_raw(p, np.int32)
    with pytest.raises(M.MlemError):
        M.read_raw(str(tmp_path / "missing"), np.int32)
    p = str(tmp_path / "empty")
    open(p, "wb").close()
    assert M.read_raw(p, np.int32).size == 0


def test_host_copy_is_memcpy():
    """mlem_host_copy (the csr4 loader's fill of its staging buffers): byte for byte, any size and thread count,
    slices split on 4 KiB boundaries."""
    import ctypes as C
    lib = M.host_lib()
    fn = lib.mlem_host_copy
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int]
    rng = np.random.default_rng(9)
    for n in (0, 1, 4095, 4097, 1 << 20, (3 << 20) + 13):
        src = rng.integers(0, 256, n, dtype=np.uint8)
        for threads in (0, 1, 3, 16):
            dst = np.zeros(n + 8, np.uint8)
            assert fn(dst.ctypes.data, src.ctypes.data, n, threads) == M.MLEM_OK
            assert np.array_equal(dst[:n], src) and not dst[n:].any()
    assert fn(None, None, 0, 0) == M.MLEM_OK
    assert fn(None, src.ctypes.data, 16, 0) != M.MLEM_OK
