"""ctypes binding of the CPU oracle (oracle/liboracle.so) and of the
reference's own sptrans compiled into oracle/_ref/.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to check results against.  The engine never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(ORACLE_DIR, "liboracle.so")
REF_LIB = os.path.join(ORACLE_DIR, "_ref", "libsptrans_ref.so")
REF_SRC = "/root/reference/sptrans.h"

vp = C.c_void_p
u32, u64, i64, f32, f64, cint = C.c_uint32, C.c_uint64, C.c_int64, C.c_float, C.c_double, C.c_int

_lib = None
_ref = None


def build(ref: bool = True):
    """Compile liboracle.so (and _ref/ when the reference checkout exists)."""
    targets = ["all"] + (["ref"] if ref and os.path.exists(REF_SRC) else [])
    subprocess.run(["make", "-s", "-C", ORACLE_DIR] + targets, check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build(ref=False)
        L = C.CDLL(LIB)
        sig = {
            "oracle_csr_from_csr4": (None, [u32, vp, vp, vp, vp, vp]),
            "oracle_colsum_f64": (None, [vp, vp, vp, u32, u32, vp]),
            "oracle_colsum_f32": (None, [vp, vp, vp, u32, u32, vp]),
            "oracle_initial_image": (f32, [vp, u32, vp, u32]),
            "oracle_sptrans": (cint, [u32, u32, u64, vp, vp, vp, vp, vp, vp, u32]),
            "oracle_partition": (None, [vp, u64, u32, u32, vp, vp, vp, vp]),
            "oracle_merge_coord": (None, [vp, u32, i64, C.POINTER(i64), C.POINTER(i64)]),
            "oracle_merge_spmv_f64": (None, [vp, vp, vp, vp, u32, u64, u32, vp]),
            "oracle_spmv_f64": (None, [vp, vp, vp, vp, u32, vp]),
            "oracle_mlem_f64": (None, [vp, vp, vp, u32, u32, vp, vp, vp, u32]),
            "oracle_loglik_f64": (f64, [vp, vp, vp, u32, vp, vp]),
            "oracle_mlem_f32_omp": (f64, [vp, vp, vp, vp, vp, vp, u32, u32, vp, vp, vp, u32, cint]),
            "oracle_mlem_f64_omp": (cint, [vp, vp, vp, vp, vp, vp, u32, u32, u32, vp, vp, vp, u32, cint]),
            "oracle_max_threads": (cint, []),
        }
        for n, (r, a) in sig.items():
            fn = getattr(L, n)
            fn.restype, fn.argtypes = r, a
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref_lib() -> C.CDLL:
    global _ref
    if _ref is None:
        L = C.CDLL(REF_LIB)
        L.ref_sptrans.restype = cint
        L.ref_sptrans.argtypes = [u32, u32, u64, vp, vp, vp, vp, vp, vp, cint]
        _ref = L
    return _ref


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(vp)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def csr_from_csr4(rows, row_ends, aos_bytes):
    row_ends = _c(row_ends, np.uint64)
    aos = _c(np.frombuffer(aos_bytes, np.uint8), np.uint8)
    nnz = int(row_ends[-1]) if rows else 0
    rowptr = np.zeros(rows + 1, np.uint64)
    col = np.empty(nnz, np.uint32)
    val = np.empty(nnz, np.float32)
    lib().oracle_csr_from_csr4(rows, _p(row_ends), _p(aos), _p(rowptr), _p(col), _p(val))
    return rowptr, col, val


def colsum_f64(rowptr, col, val, rows, cols):
    out = np.zeros(cols, np.float64)
    lib().oracle_colsum_f64(_p(_c(rowptr, np.uint64)), _p(_c(col, np.uint32)), _p(_c(val, np.float32)), rows, cols,
                            _p(out))
    return out


def colsum_f32(rowptr, col, val, rows, cols):
    out = np.zeros(cols, np.float32)
    lib().oracle_colsum_f32(_p(_c(rowptr, np.uint64)), _p(_c(col, np.uint32)), _p(_c(val, np.float32)), rows, cols,
                            _p(out))
    return out


def initial_image(g, rows, norm):
    return lib().oracle_initial_image(_p(_c(g, np.int32)), rows, _p(_c(norm, np.float32)), len(norm))


def sptrans(rows, cols, rowptr, col, val, procs=1):
    rowptr = _c(rowptr, np.uint64)
    nnz = int(rowptr[-1])
    colptr = np.zeros(cols + 1, np.uint64)
    rowT = np.zeros(nnz, np.uint32)
    valT = np.zeros(nnz, np.float32) if val is not None else None
    st = lib().oracle_sptrans(rows, cols, nnz, _p(rowptr), _p(_c(col, np.uint32)),
                              _p(_c(val, np.float32)) if val is not None else None, _p(rowT), _p(colptr), _p(valT),
                              procs)
    assert st == 0
    return colptr, rowT, valT


def ref_sptrans(rows, cols, rowptr, col, val, threads=1):
    """The reference's own sptrans_scanTrans_specialized (oracle/_ref)."""
    rowptr = _c(rowptr, np.uint64).copy()
    nnz = int(rowptr[-1])
    col = _c(col, np.uint32).copy()
    val = _c(val, np.float32).copy()
    colptr = np.zeros(cols + 1, np.uint64)
    rowT = np.zeros(nnz, np.uint32)
    valT = np.zeros(nnz, np.float32)
    ref_lib().ref_sptrans(rows, cols, nnz, _p(rowptr), _p(col), _p(val), _p(rowT), _p(colptr), _p(valT), threads)
    return colptr, rowT, valT


def partition(rowptr, nnz, rows, D):
    seg = np.zeros(D + 1, np.uint32)
    srows = np.zeros(D, np.uint32)
    snnz = np.zeros(D, np.uint32)
    off = np.zeros(D, np.uint64)
    lib().oracle_partition(_p(_c(rowptr, np.uint64)), nnz, rows, D, _p(seg), _p(srows), _p(snnz), _p(off))
    return seg, srows, snnz, off


def merge_coord(rowptr, rows, d):
    i, j = i64(), i64()
    lib().oracle_merge_coord(_p(_c(rowptr, np.uint64)), rows, d, C.byref(i), C.byref(j))
    return i.value, j.value


def merge_spmv_f64(rowptr, col, val, x, rows, secsize):
    rowptr = _c(rowptr, np.uint64)
    y = np.zeros(rows, np.float64)
    lib().oracle_merge_spmv_f64(_p(rowptr), _p(_c(col, np.uint32)), _p(_c(val, np.float32)), _p(_c(x, np.float32)),
                                rows, int(rowptr[-1]), secsize, _p(y))
    return y


def spmv_f64(rowptr, col, val, x, rows):
    y = np.zeros(rows, np.float64)
    lib().oracle_spmv_f64(_p(_c(rowptr, np.uint64)), _p(_c(col, np.uint32)), _p(_c(val, np.float32)),
                          _p(_c(x, np.float64)), rows, _p(y))
    return y


def mlem_f64(rowptr, col, val, rows, cols, g, sens, f0, iters):
    f = _c(f0, np.float64).copy()
    lib().oracle_mlem_f64(_p(_c(rowptr, np.uint64)), _p(_c(col, np.uint32)), _p(_c(val, np.float32)), rows, cols,
                          _p(_c(g, np.int32)), _p(_c(sens, np.float64)), _p(f), iters)
    return f


def mlem_f64_omp(A, AT, g, sens, f0, iters, threads=0):
    """oracle_mlem_f64 for one image (g 1-D, f0 1-D) or B images (g [B, rows], f0 [B, cols]) with OpenMP, FP over
    rows of A and BP over rows of AT (the stable transpose of A); bitwise equal to mlem_f64 image by image.
    Returns f64 images shaped like f0."""
    g = np.asarray(g)
    f0 = np.asarray(f0)
    single = g.ndim == 1
    G = _c(g.reshape(1, -1) if single else g, np.int32)
    F0 = f0.reshape(1, -1) if single else f0
    B = G.shape[0]
    if F0.shape != (B, A.cols) or G.shape[1] != A.rows:
        raise ValueError("mlem_f64_omp: counts or images of the wrong shape")
    Gi = np.ascontiguousarray(G.T)  # [rows][B]
    F = np.ascontiguousarray(F0.astype(np.float64).T)  # [cols][B]
    st = lib().oracle_mlem_f64_omp(_p(A.rowptr), _p(A.col), _p(A.val), _p(AT.rowptr), _p(AT.col), _p(AT.val),
                                   A.rows, A.cols, B, _p(Gi), _p(_c(sens, np.float64)), _p(F), iters, threads)
    if st != 0:
        raise ValueError(f"oracle_mlem_f64_omp failed ({st})")
    out = np.ascontiguousarray(F.T)
    return out[0] if single else out


def loglik_f64(rowptr, col, val, rows, g, f):
    return lib().oracle_loglik_f64(_p(_c(rowptr, np.uint64)), _p(_c(col, np.uint32)), _p(_c(val, np.float32)), rows,
                                   _p(_c(g, np.int32)), _p(_c(f, np.float64)))


def mlem_f32_omp(A, AT, g, sens, f0, iters, threads=0):
    """FP32 OpenMP CPU MLEM (the CPU baseline); returns (image, seconds)."""
    f = _c(f0, np.float32).copy()
    secs = lib().oracle_mlem_f32_omp(_p(A.rowptr), _p(A.col), _p(A.val), _p(AT.rowptr), _p(AT.col), _p(AT.val),
                                     A.rows, A.cols, _p(_c(g, np.int32)), _p(_c(sens, np.float32)), _p(f), iters,
                                     threads)
    return f, secs


def max_threads() -> int:
    return lib().oracle_max_threads()


# ------------------------------------------------------------ 16-bit values
# Restatement of this engine's 16-bit matrix values (the reference stores f32
# only; its precision extension is the per-row scale of the engine's
# documentation): scale_r = max_k |a_rk| (1 for an all-zero row) times 2^-14
# for fp16, 1 otherwise (f32 product); t = a * (1 / scale_r) in f32; the code
# is t rounded to nearest even in the type; the value the kernels use is
# scale_r * decode(code).  Pinned by known codes and ties
# (tests/test_oracle.py::test_quantize_values_known_codes).

def bf16_code(t):
    u = np.asarray(t, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def bf16_value(c):
    return (np.asarray(c, np.uint32) << 16).astype(np.uint32).view(np.float32)


def uf16_code(t, exp_bits=4):
    """Unsigned 16-bit float: exp_bits exponent and 16 - exp_bits mantissa bits, t in [0, 1] (clamped), the top
    binade holds 1.0, code 0 is the smallest value 2^-(2^exp_bits - 1) (no zero)."""
    sh, emax = 7 + exp_bits, (1 << exp_bits) - 1
    t = np.clip(np.asarray(t, np.float32), 0.0, 1.0).astype(np.float32)
    u = t.view(np.uint32).astype(np.int64)
    ur = u + ((1 << (sh - 1)) - 1) + ((u >> sh) & 1)
    c = (ur >> sh) - ((127 - emax) << (16 - exp_bits))
    return np.where(c < 0, 0, c).astype(np.uint16)


def uf16_value(c, exp_bits=4):
    sh, emax = 7 + exp_bits, (1 << exp_bits) - 1
    return ((np.asarray(c, np.uint32) << sh) + ((127 - emax) << 23)).astype(np.uint32).view(np.float32)


def quantize_values(rowptr, val, rows, kind):
    """The values the 16-bit kernels multiply with, as f64 (scale_r * decode(code)): kind 1 bf16, 2 fp16, 3 uf16."""
    rowptr = _c(rowptr, np.uint64).astype(np.int64)
    val = _c(val, np.float32)
    lens = np.diff(rowptr)
    row_of = np.repeat(np.arange(rows), lens)
    amax = np.zeros(rows, np.float32)
    np.maximum.at(amax, row_of, np.abs(val))
    mult = np.float32(2.0 ** -14) if kind == 2 else np.float32(1.0)
    scale = (np.where(amax > 0, amax, np.float32(1.0)).astype(np.float32) * mult).astype(np.float32)
    inv = (np.float32(1.0) / scale).astype(np.float32)
    t = (val * inv[row_of]).astype(np.float32)
    if kind == 1:
        dec = bf16_value(bf16_code(t))
    elif kind == 2:
        dec = t.astype(np.float16).astype(np.float32)
    elif kind == 3:
        dec = uf16_value(uf16_code(t))
    else:
        raise ValueError(kind)
    return scale[row_of].astype(np.float64) * dec.astype(np.float64)
