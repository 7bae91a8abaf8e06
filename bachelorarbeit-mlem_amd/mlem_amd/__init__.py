"""ctypes bindings of the MLEM engine's C ABI (include/mlem_host.h, include/mlem_hip.h).

This module is glue for tests, the benchmark and scripting: every computation
runs in libmlem_host.so (CPU preprocessing, formats, generator) or
libmlem_hip.so (gfx950 kernels).  Nothing here computes a result itself, and
the device half raises MlemError when libmlem_hip.so is missing: there is no
fallback path.

Reference interfaces mirrored (file:line in the reference repository):
  * Csr4Matrix / csr_format_for_cuda / Vector<int>  (csr4matrix.cpp:36-254, main.cu:518-554, vector.hpp:33-53)
  * calcColumnSums, init, partitionMatrix           (main.cu:558-611, 83-97)
  * sptrans_scanTrans_specialized                   (sptrans.h:397-545)
  * the seven kernels of kernel.cuh:7-19 and mlem() (main.cu:167-514)
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass
from typing import Optional

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_ROOT, "lib")
BIN_DIR = os.path.join(PKG_ROOT, "bin")
# MLEM_HOST_LIB / MLEM_CLI select other builds of the host library and the
# command line (the sanitizer build of `make asan`).
HOST_LIB_PATH = os.environ.get("MLEM_HOST_LIB") or os.path.join(LIB_DIR, "libmlem_host.so")
HIP_LIB_PATH = os.path.join(LIB_DIR, "libmlem_hip.so")
CLI_PATH = os.environ.get("MLEM_CLI") or os.path.join(BIN_DIR, "mlem")

MLEM_OK = 0
STATUS_NAMES = {
    0: "MLEM_OK", 1: "MLEM_ERR_INVALID_ARGUMENT", 2: "MLEM_ERR_HIP", 3: "MLEM_ERR_OUT_OF_MEMORY",
    4: "MLEM_ERR_UNSUPPORTED", 5: "MLEM_ERR_NO_DEVICE", 6: "MLEM_ERR_IO", 7: "MLEM_ERR_FORMAT",
}


class MlemError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


vp = C.c_void_p
u8, u32, u64, i32, i64, f32, f64 = C.c_uint8, C.c_uint32, C.c_uint64, C.c_int32, C.c_int64, C.c_float, C.c_double
sz = C.c_size_t
cint = C.c_int


class Scanner(C.Structure):
    _fields_ = [("n_layers", u32), ("layer", (f32 * 5) * 16), ("n_blocks", u32), ("blocksize", u32),
                ("block_gap", f32 * 255), ("n_rings", u32), ("ring_gap", f32 * 255),
                ("voxel_size", f32 * 3), ("grid", u32 * 3)]


class Geometry(C.Structure):
    _fields_ = [("n_rings", u32), ("n_per_ring", u32), ("radius", f32), ("crystal", f32 * 3),
                ("ring_pitch", f32), ("grid", u32 * 3), ("voxel", f32 * 3), ("sub_rays", u32)]

    def __repr__(self):
        return (f"Geometry(rings={self.n_rings}, per_ring={self.n_per_ring}, radius={self.radius}, "
                f"crystal={list(self.crystal)}, pitch={self.ring_pitch}, grid={list(self.grid)}, "
                f"voxel={list(self.voxel)}, sub_rays={self.sub_rays})")


class CsrHost(C.Structure):
    _fields_ = [("rows", u32), ("cols", u32), ("nnz", u64), ("rowptr", C.POINTER(u64)),
                ("col", C.POINTER(u32)), ("val", C.POINTER(f32))]


class Opts(C.Structure):
    _fields_ = [("spmv", u32), ("using_trans", u32), ("secsize_fp", u32), ("secsize_bp", u32),
                ("fp_lanes", u32), ("bp_lanes", u32), ("profile", u32), ("use_graph", u32),
                ("bp_spmv", u32), ("tile_width", u32), ("tile_rows", u32), ("tile_groups", u32), ("values", u32)]


def geometry(n_rings, n_per_ring, radius, crystal, ring_pitch, grid, voxel, sub_rays) -> Geometry:
    g = Geometry()
    g.n_rings, g.n_per_ring, g.radius = n_rings, n_per_ring, radius
    g.crystal[:] = list(crystal)
    g.ring_pitch = ring_pitch
    g.grid[:] = list(grid)
    g.voxel[:] = list(voxel)
    g.sub_rays = sub_rays
    return g


# Scanner presets.  RING2D_288: 288 point crystals on a 40 mm ring, 64x64
# voxels of 1 mm, exact Siddon lengths (41,328 LOR rows).  RING3D_1152: two
# axial rings of 576 crystals (1 x 10 x 10 mm, inner radius 35.5 mm, ring
# pitch 10 mm), 140x140x40 voxels of 0.5 mm, 22 sub-rays per LOR (662,976
# rows x 784,000 columns, 1,527,217,536 nonzeros: 12.2 GB as f32 + u32; a
# sub-ray crosses 124 voxels on average and the sub-rays of one LOR share
# about 10% of them, so the count grows with the sub-rays: 16 give 1.15e9,
# and crystal size, radius or voxel size move it by a few percent only).
# RING3D_1152_FINE: the same scanner, 280x280x80
# voxels of 0.25 mm and 128 sub-rays.  RING3D_576_MID: two rings of 288
# (2 x 9 x 10 mm, radius 35.5 mm), 96x96x24 voxels of 0.75 mm, 16 sub-rays
# (165,600 rows x 221,184 columns).
def RING2D_288() -> Geometry:
    return geometry(1, 288, 40.0, (0.0, 0.0, 0.0), 0.0, (64, 64, 1), (1.0, 1.0, 1.0), 0)


def RING3D_1152() -> Geometry:
    return geometry(2, 576, 35.5, (1.0, 10.0, 10.0), 10.0, (140, 140, 40), (0.5, 0.5, 0.5), 22)


def RING3D_1152_FINE() -> Geometry:
    return geometry(2, 576, 35.5, (1.0, 10.0, 10.0), 10.0, (280, 280, 80), (0.25, 0.25, 0.25), 128)


def RING3D_576_MID() -> Geometry:
    return geometry(2, 288, 35.5, (2.0, 9.0, 10.0), 9.0, (96, 96, 24), (0.75, 0.75, 0.75), 16)


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(vp)


def _load(path: str, what: str) -> C.CDLL:
    if not os.path.exists(path):
        raise MlemError(4, f"{what} not built: {path} is missing (run __graft_entry__.build())")
    return C.CDLL(path)


_host = None
_hip = None


def host_lib() -> C.CDLL:
    global _host
    if _host is None:
        lib = _load(HOST_LIB_PATH, "libmlem_host.so")
        sigs = {
            "mlem_host_last_error": (C.c_char_p, []),
            "mlem_host_status_string": (C.c_char_p, [cint]),
            "mlem_csr4_open": (cint, [C.c_char_p, C.POINTER(vp)]),
            "mlem_csr4_close": (cint, [vp]),
            "mlem_csr4_info": (cint, [vp, C.POINTER(u32), C.POINTER(u32), C.POINTER(u64), C.POINTER(u8), C.POINTER(u32)]),
            "mlem_csr4_scanner": (cint, [vp, C.POINTER(Scanner)]),
            "mlem_csr4_row_ends": (C.POINTER(u64), [vp]),
            "mlem_csr4_data": (vp, [vp]),
            "mlem_csr4_elements_in_row": (u64, [vp, u32]),
            "mlem_csr4_to_csr": (cint, [vp, vp, vp, vp, cint]),
            "mlem_csr4_write": (cint, [C.c_char_p, u32, u32, u64, vp, vp, vp, C.POINTER(Scanner), u8]),
            "mlem_read_raw": (cint, [C.c_char_p, sz, C.POINTER(vp), C.POINTER(u64)]),
            "mlem_write_raw": (cint, [C.c_char_p, vp, sz, u64]),
            "mlem_host_free": (None, [vp]),
            "mlem_colsum_f32": (cint, [vp, vp, vp, u32, u32, vp]),
            "mlem_initial_image": (cint, [vp, u32, vp, u32, vp, C.POINTER(f32)]),
            "mlem_transpose": (cint, [u32, u32, u64, vp, vp, vp, vp, vp, vp, cint]),
            "mlem_partition": (cint, [vp, u64, u32, u32, vp, vp, vp, vp, vp]),
            "mlem_lor_order": (cint, [vp, vp, vp, u32, u32, vp, cint, vp]),
            "mlem_geometry_rows": (u32, [C.POINTER(Geometry)]),
            "mlem_geometry_cols": (u32, [C.POINTER(Geometry)]),
            "mlem_gen_system_matrix": (cint, [C.POINTER(Geometry), u64, u32, u32, cint, C.POINTER(CsrHost)]),
            "mlem_csr_host_free": (None, [C.POINTER(CsrHost)]),
            "mlem_gen_phantom": (cint, [C.POINTER(Geometry), u64, u32, f32, f32, vp]),
            "mlem_gen_sinogram": (cint, [u32, vp, vp, vp, vp, f64, u64, cint, vp, vp]),
            "mlem_poisson_sample": (i64, [f64, u64, u64]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _host = lib
    return _host


def _hcheck(st: int):
    if st != MLEM_OK:
        raise MlemError(st, host_lib().mlem_host_last_error().decode())


# --------------------------------------------------------------------------
# host-side data
# --------------------------------------------------------------------------
@dataclass
class Csr:
    """A CSR matrix on the host: rowptr u64[rows+1], col u32[nnz], val f32[nnz]."""
    rows: int
    cols: int
    rowptr: np.ndarray
    col: np.ndarray
    val: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col.astype(np.int64), self.rowptr.astype(np.int64)),
                             shape=(self.rows, self.cols))

    @staticmethod
    def from_scipy(m) -> "Csr":
        m = m.tocsr()
        return Csr(m.shape[0], m.shape[1], np.ascontiguousarray(m.indptr, dtype=np.uint64),
                   np.ascontiguousarray(m.indices, dtype=np.uint32), np.ascontiguousarray(m.data, dtype=np.float32))


class _Owned:
    """Keeps a library-allocated CSR alive while numpy views of it exist."""

    def __init__(self, st: CsrHost):
        self.st = st
        weakref.finalize(self, host_lib().mlem_csr_host_free, C.byref(st))


class _Buf:
    """Exposes library memory to numpy; numpy keeps this object (and so the owner) as the array's base."""

    def __init__(self, addr: int, n: int, dtype, owner):
        self.owner = owner
        self.__array_interface__ = {"shape": (n,), "typestr": np.dtype(dtype).str, "data": (addr, False),
                                    "version": 3}


def _view(ptr, n, dtype, owner):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.asarray(_Buf(C.cast(ptr, vp).value, n, dtype, owner))


def gen_system_matrix(geo: Geometry, seed: int = 1, row_begin: int = 0, row_end: int = 0,
                      threads: int = 0) -> Csr:
    st = CsrHost()
    _hcheck(host_lib().mlem_gen_system_matrix(C.byref(geo), seed, row_begin, row_end, threads, C.byref(st)))
    owner = _Owned(st)
    rowptr = _view(st.rowptr, st.rows + 1, np.uint64, owner)
    col = _view(st.col, st.nnz, np.uint32, owner)
    val = _view(st.val, st.nnz, np.float32, owner)
    return Csr(st.rows, st.cols, rowptr, col, val)


def geometry_rows(geo: Geometry) -> int:
    return host_lib().mlem_geometry_rows(C.byref(geo))


def geometry_cols(geo: Geometry) -> int:
    return host_lib().mlem_geometry_cols(C.byref(geo))


def gen_phantom(geo: Geometry, seed: int = 1, n_shapes: int = 5, a_min: float = 0.5, a_max: float = 4.0) -> np.ndarray:
    x = np.zeros(geometry_cols(geo), dtype=np.float32)
    _hcheck(host_lib().mlem_gen_phantom(C.byref(geo), seed, n_shapes, a_min, a_max, _ptr(x)))
    return x


def gen_sinogram(A: Csr, x: np.ndarray, total_counts: float, seed: int = 1, threads: int = 0,
                 return_expected: bool = False):
    y = np.zeros(A.rows, dtype=np.int32)
    lam = np.zeros(A.rows, dtype=np.float64) if return_expected else None
    x = np.ascontiguousarray(x, dtype=np.float32)
    _hcheck(host_lib().mlem_gen_sinogram(A.rows, _ptr(A.rowptr), _ptr(A.col), _ptr(A.val), _ptr(x),
                                         float(total_counts), seed, threads, _ptr(y), _ptr(lam)))
    return (y, lam) if return_expected else y


def poisson_sample(lam: float, seed: int, idx: int) -> int:
    return int(host_lib().mlem_poisson_sample(lam, seed, idx))


def transpose(A: Csr, threads: int = 0, values: bool = True) -> Csr:
    """Stable counting-sort transpose (sptrans_scanTrans_specialized)."""
    colptr = np.zeros(A.cols + 1, dtype=np.uint64)
    rowT = np.empty(A.nnz, dtype=np.uint32)
    valT = np.empty(A.nnz, dtype=np.float32) if values else None
    _hcheck(host_lib().mlem_transpose(A.rows, A.cols, A.nnz, _ptr(A.rowptr), _ptr(A.col), _ptr(A.val),
                                      _ptr(colptr), _ptr(rowT), _ptr(valT), threads))
    return Csr(A.cols, A.rows, colptr, rowT, valT if values else np.zeros(0, np.float32))


def partition(rowptr: np.ndarray, nnz: int, rows: int, parts: int):
    """partitionMatrix (main.cu:593-611) -> (segments, segment_rows, segment_nnzs(u32), segment_nnzs64, offsets)."""
    seg = np.zeros(parts + 1, np.uint32)
    srows = np.zeros(parts, np.uint32)
    snnz = np.zeros(parts, np.uint32)
    snnz64 = np.zeros(parts, np.uint64)
    off = np.zeros(parts, np.uint64)
    rowptr = np.ascontiguousarray(rowptr, dtype=np.uint64)
    _hcheck(host_lib().mlem_partition(_ptr(rowptr), nnz, rows, parts, _ptr(seg), _ptr(srows), _ptr(snnz),
                                      _ptr(snnz64), _ptr(off)))
    return seg, srows, snnz, snnz64, off


def lor_order(A: Csr, grid, threads: int = 0) -> np.ndarray:
    """Row order for the batched layouts (mlem_lor_order): order[i] = place of LOR i, from each row's transaxial
    line fitted to its voxels on the image grid (nx, ny, nz)."""
    g = np.ascontiguousarray(grid, dtype=np.uint32)
    order = np.empty(A.rows, np.uint32)
    _hcheck(host_lib().mlem_lor_order(_ptr(A.rowptr), _ptr(A.col), _ptr(A.val), A.rows, A.cols, _ptr(g), threads,
                                      _ptr(order)))
    return order


def colsum_f32(A: Csr) -> np.ndarray:
    norm = np.zeros(A.cols, np.float32)
    _hcheck(host_lib().mlem_colsum_f32(_ptr(A.rowptr), _ptr(A.col), _ptr(A.val), A.rows, A.cols, _ptr(norm)))
    return norm


def initial_image(g: np.ndarray, rows: int, norm: np.ndarray):
    f0 = np.zeros(len(norm), np.float32)
    init = f32()
    g = np.ascontiguousarray(g, dtype=np.int32)
    norm = np.ascontiguousarray(norm, dtype=np.float32)
    _hcheck(host_lib().mlem_initial_image(_ptr(g), rows, _ptr(norm), len(norm), _ptr(f0), C.byref(init)))
    return f0, init.value


class Csr4File:
    """An open csr4 file (Csr4Matrix)."""

    def __init__(self, path: str):
        self._h = vp()
        _hcheck(host_lib().mlem_csr4_open(path.encode(), C.byref(self._h)))
        r, c, n, fl, scb = u32(), u32(), u64(), u8(), u32()
        _hcheck(host_lib().mlem_csr4_info(self._h, C.byref(r), C.byref(c), C.byref(n), C.byref(fl), C.byref(scb)))
        self.rows, self.cols, self.nnz, self.flags, self.scan_config_bytes = r.value, c.value, n.value, fl.value, scb.value

    def scanner(self) -> Scanner:
        s = Scanner()
        _hcheck(host_lib().mlem_csr4_scanner(self._h, C.byref(s)))
        return s

    def row_ends(self) -> np.ndarray:
        p = host_lib().mlem_csr4_row_ends(self._h)
        return np.ctypeslib.as_array(p, shape=(self.rows,)).copy() if self.rows else np.zeros(0, np.uint64)

    def elements_in_row(self, r: int) -> int:
        return int(host_lib().mlem_csr4_elements_in_row(self._h, r))

    def to_csr(self, threads: int = 0) -> Csr:
        rowptr = np.zeros(self.rows + 1, np.uint64)
        col = np.empty(self.nnz, np.uint32)
        val = np.empty(self.nnz, np.float32)
        _hcheck(host_lib().mlem_csr4_to_csr(self._h, _ptr(rowptr), _ptr(col), _ptr(val), threads))
        return Csr(self.rows, self.cols, rowptr, col, val)

    def close(self):
        if self._h:
            host_lib().mlem_csr4_close(self._h)
            self._h = vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_csr4(path: str, A: Csr, scanner: Optional[Scanner] = None, flags: int = 0):
    _hcheck(host_lib().mlem_csr4_write(path.encode(), A.rows, A.cols, A.nnz, _ptr(A.rowptr), _ptr(A.col),
                                       _ptr(A.val), C.byref(scanner) if scanner is not None else None, flags))


def read_csr4(path: str, threads: int = 0) -> Csr:
    with Csr4File(path) as f:
        return f.to_csr(threads)


def read_raw(path: str, dtype) -> np.ndarray:
    dtype = np.dtype(dtype)
    p, n = vp(), u64()
    _hcheck(host_lib().mlem_read_raw(path.encode(), dtype.itemsize, C.byref(p), C.byref(n)))
    try:
        if n.value == 0:
            return np.zeros(0, dtype)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(n.value * dtype.itemsize,)).view(dtype).copy()
    finally:
        host_lib().mlem_host_free(p)


def write_raw(path: str, a: np.ndarray):
    a = np.ascontiguousarray(a)
    _hcheck(host_lib().mlem_write_raw(path.encode(), _ptr(a), a.dtype.itemsize, a.size))


from .device import *  # noqa: E402,F401,F403  (device half: libmlem_hip.so)
