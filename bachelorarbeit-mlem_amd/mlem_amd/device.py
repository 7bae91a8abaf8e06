"""Device half of the ctypes bindings: libmlem_hip.so (include/mlem_hip.h).

Importing this module does not touch the GPU; the library is loaded on first
use and MlemError is raised if it is not built.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import (HIP_LIB_PATH, MLEM_OK, Csr, MlemError, Opts, _load, _ptr, cint, f32, f64, host_lib, i32, sz, u32,
               u64, vp)

__all__ = ["hip_lib", "device_count", "device_info", "DeviceArray", "Stream", "Event", "ResidentMatrix", "Recon",
           "Batch", "BatchOpts", "Server",
           "mlem_run", "default_opts", "opts", "merge_workspace_bytes", "fp_merge", "fp_merge_noalloc", "bp_merge",
           "fp_csr_vector",
           "bp_csr_vector", "bp_scatter", "fp_ratio_merge", "fp_ratio_csr_vector", "bp_update_merge",
           "bp_update_csr_vector", "ratio", "update", "colsum_f64", "colsum_scatter_f64", "fill_f32", "copy_f32",
           "device_sync",
           "version"]

_hip = None


def hip_lib() -> C.CDLL:
    global _hip
    if _hip is None:
        host_lib()  # first, so that libmlem_hip.so binds to the host library selected (by its soname)
        lib = _load(HIP_LIB_PATH, "libmlem_hip.so")
        P = C.POINTER
        sigs = {
            "mlem_last_error": (C.c_char_p, []),
            "mlem_status_string": (C.c_char_p, [cint]),
            "mlem_version": (C.c_char_p, []),
            "mlem_device_count": (cint, [P(cint)]),
            "mlem_set_device": (cint, [cint]),
            "mlem_device_info": (cint, [cint, C.c_char_p, sz, P(u64), P(cint)]),
            "mlem_device_alloc": (cint, [P(vp), sz]),
            "mlem_device_free": (cint, [vp]),
            "mlem_memcpy_h2d": (cint, [vp, vp, sz, vp]),
            "mlem_memcpy_d2h": (cint, [vp, vp, sz, vp]),
            "mlem_memcpy_d2d": (cint, [vp, vp, sz, vp]),
            "mlem_stream_create": (cint, [P(vp)]),
            "mlem_stream_destroy": (cint, [vp]),
            "mlem_stream_sync": (cint, [vp]),
            "mlem_device_sync": (cint, []),
            "mlem_event_create": (cint, [P(vp)]),
            "mlem_event_destroy": (cint, [vp]),
            "mlem_event_record": (cint, [vp, vp]),
            "mlem_event_elapsed_ms": (cint, [vp, vp, P(f32)]),
            "mlem_fill_f32": (cint, [vp, f32, u64, vp]),
            "mlem_copy_f32": (cint, [vp, vp, u64, vp]),
            "mlem_ratio": (cint, [vp, vp, u32, vp]),
            "mlem_update": (cint, [vp, vp, vp, u32, vp]),
            "mlem_merge_workspace_bytes": (sz, [u32, u64, u32]),
            "mlem_fp_merge": (cint, [vp, vp, vp, vp, vp, u32, u32, u64, vp, sz, vp]),
            "mlem_fp_merge_noalloc": (cint, [vp, vp, vp, vp, vp, u32, u32, u64, vp]),
            "mlem_bp_merge": (cint, [vp, vp, vp, vp, vp, u32, u32, u64, vp, sz, vp]),
            "mlem_fp_csr_vector": (cint, [vp, vp, vp, vp, vp, u32, u64, u32, vp]),
            "mlem_bp_csr_vector": (cint, [vp, vp, vp, vp, vp, u32, u64, u32, vp]),
            "mlem_bp_scatter": (cint, [vp, vp, vp, vp, vp, u32, u32, u64, u32, vp]),
            "mlem_fp_ratio_merge": (cint, [vp, vp, vp, vp, vp, vp, u32, u32, u64, vp, sz, vp]),
            "mlem_fp_ratio_csr_vector": (cint, [vp, vp, vp, vp, vp, vp, u32, u64, u32, vp]),
            "mlem_bp_update_merge": (cint, [vp, vp, vp, vp, vp, vp, u32, u32, u64, vp, sz, vp]),
            "mlem_bp_update_csr_vector": (cint, [vp, vp, vp, vp, vp, vp, u32, u64, u32, vp]),
            "mlem_colsum_f64": (cint, [vp, vp, u32, u64, vp, vp, vp]),
            "mlem_colsum_scatter_f64": (cint, [vp, vp, vp, u32, u32, u64, vp, vp, vp]),
            "mlem_opts_default": (None, [P(Opts)]),
            "mlem_matrix_create": (cint, [cint, u32, u32, u64, vp, vp, vp, vp, vp, vp, P(vp)]),
            "mlem_matrix_build_transpose": (cint, [vp]),
            "mlem_matrix_build_transpose_limited": (cint, [vp, u64, P(u32)]),
            "mlem_matrix_set_row_order": (cint, [vp, vp]),
            "mlem_matrix_create_tiled": (cint, [cint, u32, u32, u64, vp, vp, vp, u32, u32, u32, u32, u64, P(vp)]),
            "mlem_autotune": (cint, [vp, u32, u32, vp]),
            "mlem_matrix_load_csr4": (cint, [cint, C.c_char_p, cint, P(vp), vp]),
            "mlem_batch_opts_default": (None, [vp]),
            "mlem_batch_create": (cint, [vp, u32, vp, P(vp)]),
            "mlem_batch_set": (cint, [vp, u32, vp, vp]),
            "mlem_batch_deactivate": (cint, [vp, u32]),
            "mlem_batch_iterate": (cint, [vp, u32]),
            "mlem_batch_sync": (cint, [vp]),
            "mlem_batch_get_image": (cint, [vp, u32, vp]),
            "mlem_batch_info": (cint, [vp, P(u32), P(u32), P(u32), P(u64)]),
            "mlem_batch_stream": (vp, [vp]),
            "mlem_batch_destroy": (cint, [vp]),
            "mlem_server_create": (cint, [u64, vp, P(vp)]),
            "mlem_server_add_matrix": (cint, [vp, vp, u32, P(u32)]),
            "mlem_server_submit": (cint, [vp, u32, vp, vp, u32, P(u64)]),
            "mlem_server_step": (cint, [vp, u32, P(u32), P(u32)]),
            "mlem_server_poll": (cint, [vp, u64, vp, P(f64), P(cint)]),
            "mlem_server_info": (cint, [vp, P(u64), P(u64), P(u32), P(u64)]),
            "mlem_server_stats": (cint, [vp, P(u64), P(u64)]),
            "mlem_server_destroy": (cint, [vp]),
            "mlem_matrix_destroy": (cint, [vp]),
            "mlem_matrix_info": (cint, [vp, P(u32), P(u32), P(u64), P(cint), P(u64)]),
            "mlem_matrix_device_arrays": (cint, [vp, P(vp), P(vp), P(vp), P(vp), P(vp), P(vp)]),
            "mlem_matrix_download_transpose": (cint, [vp, vp, vp, vp]),
            "mlem_matrix_sensitivity": (cint, [vp, vp, vp]),
            "mlem_recon_create": (cint, [vp, vp, vp, vp, P(Opts), P(vp)]),
            "mlem_recon_iterate": (cint, [vp, u32]),
            "mlem_recon_sync": (cint, [vp]),
            "mlem_recon_get_image": (cint, [vp, vp]),
            "mlem_recon_set_image": (cint, [vp, vp]),
            "mlem_recon_kernel_times": (cint, [vp, P(f64), P(f64), P(u32)]),
            "mlem_recon_phase_times": (cint, [vp, P(f64), P(u32)]),
            "mlem_recon_reset_times": (cint, [vp]),
            "mlem_recon_loglik": (cint, [vp, P(f64)]),
            "mlem_recon_stream": (vp, [vp]),
            "mlem_recon_destroy": (cint, [vp]),
            "mlem_run": (cint, [vp, vp, vp, vp, vp, u32, P(Opts)]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _hip = lib
    return _hip


def _check(st: int):
    if st != MLEM_OK:
        raise MlemError(st, hip_lib().mlem_last_error().decode())


def version() -> str:
    return hip_lib().mlem_version().decode()


def device_count() -> int:
    n = cint()
    _check(hip_lib().mlem_device_count(C.byref(n)))
    return n.value


def device_info(dev: int = 0):
    name = C.create_string_buffer(256)
    mem, cus = u64(), cint()
    _check(hip_lib().mlem_device_info(dev, name, 256, C.byref(mem), C.byref(cus)))
    return name.value.decode(), mem.value, cus.value


def device_sync():
    _check(hip_lib().mlem_device_sync())


class Stream:
    def __init__(self):
        self.h = vp()
        _check(hip_lib().mlem_stream_create(C.byref(self.h)))

    def sync(self):
        _check(hip_lib().mlem_stream_sync(self.h))

    def close(self):
        if self.h:
            hip_lib().mlem_stream_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        self.h = vp()
        _check(hip_lib().mlem_event_create(C.byref(self.h)))

    def record(self, stream=None):
        _check(hip_lib().mlem_event_record(self.h, _sh(stream)))

    def elapsed_ms(self, later: "Event") -> float:
        ms = f32()
        _check(hip_lib().mlem_event_elapsed_ms(self.h, later.h, C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            if self.h:
                hip_lib().mlem_event_destroy(self.h)
        except Exception:
            pass


def _sh(stream):
    if stream is None:
        return None
    if isinstance(stream, Stream):
        return stream.h
    return stream


class DeviceArray:
    """A 1-D device buffer of a numpy dtype."""

    def __init__(self, n: int, dtype):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        self.h = vp()
        _check(hip_lib().mlem_device_alloc(C.byref(self.h), self.nbytes))

    @staticmethod
    def from_numpy(a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = DeviceArray(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray, stream=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size != self.n:
            raise ValueError("size mismatch")
        _check(hip_lib().mlem_memcpy_h2d(self.h, _ptr(a), self.nbytes, _sh(stream)))

    def numpy(self, stream=None) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        _check(hip_lib().mlem_memcpy_d2h(_ptr(out), self.h, self.nbytes, _sh(stream)))
        return out

    @property
    def ptr(self):
        return self.h

    def free(self):
        if self.h:
            hip_lib().mlem_device_free(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _dp(x):
    if x is None:
        return None
    if isinstance(x, DeviceArray):
        return x.h
    return x


def default_opts() -> Opts:
    o = Opts()
    hip_lib().mlem_opts_default(C.byref(o))
    return o


def opts(spmv=0, using_trans=0, secsize_fp=8, secsize_bp=8, fp_lanes=0, bp_lanes=0, profile=0, use_graph=0,
         bp_spmv=0, tile_width=0, tile_rows=0, tile_groups=0, values=0) -> Opts:
    """bp_spmv: 0 = BP uses the spmv algorithm too, 1 = merge-based BP, 2 = CSR-vector BP.
    tile_width: 0 = automatic LDS tiling of the merge-based kernels, 1 = plain merge path, else the segment width.
    tile_rows: rows per tile block (0 = default); tile_groups: workgroups per blocked SpMV (0 = default).
    values: stored matrix values of the blocked kernels, 0 = f32, 1 = bf16, 2 = fp16."""
    o = Opts()
    o.bp_spmv = bp_spmv
    o.tile_width = tile_width
    o.tile_rows = tile_rows
    o.tile_groups = tile_groups
    o.values = values
    o.spmv, o.using_trans, o.secsize_fp, o.secsize_bp = spmv, using_trans, secsize_fp, secsize_bp
    o.fp_lanes, o.bp_lanes, o.profile, o.use_graph = fp_lanes, bp_lanes, profile, use_graph
    return o


# ---------------------------------------------------------------- kernels
def fill_f32(dst: DeviceArray, value: float, stream=None):
    _check(hip_lib().mlem_fill_f32(_dp(dst), value, dst.n, _sh(stream)))


def copy_f32(dst: DeviceArray, src: DeviceArray, stream=None):
    """dst[:src.n] = src with the streaming copy kernel."""
    if dst.n < src.n:
        raise MlemError(1, "copy_f32: destination too small")
    _check(hip_lib().mlem_copy_f32(_dp(dst), _dp(src), src.n, _sh(stream)))


def ratio(g, fwproj, rows, stream=None):
    _check(hip_lib().mlem_ratio(_dp(g), _dp(fwproj), rows, _sh(stream)))


def update(f, norm, bwproj, cols, stream=None):
    _check(hip_lib().mlem_update(_dp(f), _dp(norm), _dp(bwproj), cols, _sh(stream)))


def merge_workspace_bytes(rows: int, nnz: int, secsize: int) -> int:
    return int(hip_lib().mlem_merge_workspace_bytes(rows, nnz, secsize))


def fp_merge(rowptr, col, val, f, out, secsize, rows, nnz, ws, ws_bytes, stream=None):
    _check(hip_lib().mlem_fp_merge(_dp(rowptr), _dp(col), _dp(val), _dp(f), _dp(out), secsize, rows, nnz, _dp(ws),
                                   ws_bytes, _sh(stream)))


def fp_merge_noalloc(rowptr, col, val, f, out, secsize, rows, nnz, stream=None):
    """fp_merge with the library's fixed scratch instead of a workspace (allocates nothing)."""
    _check(hip_lib().mlem_fp_merge_noalloc(_dp(rowptr), _dp(col), _dp(val), _dp(f), _dp(out), secsize, rows, nnz,
                                           _sh(stream)))


def bp_merge(rowptrT, colT, valT, correl, out, secsize, cols, nnz, ws, ws_bytes, stream=None):
    _check(hip_lib().mlem_bp_merge(_dp(rowptrT), _dp(colT), _dp(valT), _dp(correl), _dp(out), secsize, cols, nnz,
                                   _dp(ws), ws_bytes, _sh(stream)))


def fp_csr_vector(rowptr, col, val, f, out, rows, nnz, lanes=0, stream=None):
    _check(hip_lib().mlem_fp_csr_vector(_dp(rowptr), _dp(col), _dp(val), _dp(f), _dp(out), rows, nnz, lanes,
                                        _sh(stream)))


def bp_csr_vector(rowptrT, colT, valT, correl, out, cols, nnz, lanes=0, stream=None):
    _check(hip_lib().mlem_bp_csr_vector(_dp(rowptrT), _dp(colT), _dp(valT), _dp(correl), _dp(out), cols, nnz, lanes,
                                        _sh(stream)))


def bp_scatter(rowptr, col, val, correl, out, rows, cols, nnz, lanes=0, stream=None):
    _check(hip_lib().mlem_bp_scatter(_dp(rowptr), _dp(col), _dp(val), _dp(correl), _dp(out), rows, cols, nnz, lanes,
                                     _sh(stream)))


def fp_ratio_merge(rowptr, col, val, f, g, out, secsize, rows, nnz, ws, ws_bytes, stream=None):
    _check(hip_lib().mlem_fp_ratio_merge(_dp(rowptr), _dp(col), _dp(val), _dp(f), _dp(g), _dp(out), secsize, rows,
                                         nnz, _dp(ws), ws_bytes, _sh(stream)))


def fp_ratio_csr_vector(rowptr, col, val, f, g, out, rows, nnz, lanes=0, stream=None):
    _check(hip_lib().mlem_fp_ratio_csr_vector(_dp(rowptr), _dp(col), _dp(val), _dp(f), _dp(g), _dp(out), rows, nnz,
                                              lanes, _sh(stream)))


def bp_update_merge(rowptrT, colT, valT, ratio_, norm, f, secsize, cols, nnz, ws, ws_bytes, stream=None):
    _check(hip_lib().mlem_bp_update_merge(_dp(rowptrT), _dp(colT), _dp(valT), _dp(ratio_), _dp(norm), _dp(f),
                                          secsize, cols, nnz, _dp(ws), ws_bytes, _sh(stream)))


def bp_update_csr_vector(rowptrT, colT, valT, ratio_, norm, f, cols, nnz, lanes=0, stream=None):
    _check(hip_lib().mlem_bp_update_csr_vector(_dp(rowptrT), _dp(colT), _dp(valT), _dp(ratio_), _dp(norm), _dp(f),
                                               cols, nnz, lanes, _sh(stream)))


def colsum_f64(rowptrT, valT, cols, nnz, sens64=None, sens32=None, stream=None):
    _check(hip_lib().mlem_colsum_f64(_dp(rowptrT), _dp(valT), cols, nnz, _dp(sens64), _dp(sens32), _sh(stream)))


def colsum_scatter_f64(rowptr, col, val, rows, cols, nnz, sens64=None, sens32=None, stream=None):
    _check(hip_lib().mlem_colsum_scatter_f64(_dp(rowptr), _dp(col), _dp(val), rows, cols, nnz, _dp(sens64),
                                             _dp(sens32), _sh(stream)))


# ---------------------------------------------------------------- handles
class BatchOpts(C.Structure):
    _fields_ = [("values", u32), ("per_pass", u32), ("tile_width", u32), ("tile_rows", u32), ("tile_groups", u32),
                ("tile_threads", u32), ("use_graph", u32)]


class Batch:
    """Many reconstructions over one resident matrix (A and A^T), iterated together (mlem_batch_*)."""

    def __init__(self, matrix: "ResidentMatrix", slots: int, values: int = 0, per_pass: int = 4, tile_width: int = 0,
                 tile_rows: int = 0, tile_groups: int = 0, tile_threads: int = 0, use_graph: int = 0):
        self.h = vp()
        self.matrix = matrix
        o = BatchOpts(values, per_pass, tile_width, tile_rows, tile_groups, tile_threads, use_graph)
        _check(hip_lib().mlem_batch_create(matrix.h, slots, C.byref(o), C.byref(self.h)))
        self.slots = slots

    def set(self, slot: int, g: np.ndarray, f0: np.ndarray):
        g = np.ascontiguousarray(g, dtype=np.int32)
        f0 = np.ascontiguousarray(f0, dtype=np.float32)
        if g.size < self.matrix.rows or f0.size != self.matrix.cols:
            raise MlemError(1, "Batch.set: sinogram or image of the wrong size")
        _check(hip_lib().mlem_batch_set(self.h, slot, _ptr(g), _ptr(f0)))

    def deactivate(self, slot: int):
        _check(hip_lib().mlem_batch_deactivate(self.h, slot))

    def iterate(self, iters: int = 1):
        _check(hip_lib().mlem_batch_iterate(self.h, iters))

    def sync(self):
        _check(hip_lib().mlem_batch_sync(self.h))

    def image(self, slot: int) -> np.ndarray:
        out = np.empty(self.matrix.cols, np.float32)
        _check(hip_lib().mlem_batch_get_image(self.h, slot, _ptr(out)))
        return out

    def info(self):
        a, b, c, d = u32(), u32(), u32(), u64()
        _check(hip_lib().mlem_batch_info(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return dict(slots=a.value, per_pass=b.value, passes=c.value, device_bytes=d.value)

    def close(self):
        if self.h:
            hip_lib().mlem_batch_destroy(self.h)
            self.h = vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous batching of reconstruction requests over resident matrices (mlem_server_*).

    Each matrix added gets a batch of ``slots``; requests queue per matrix, join
    between two iterations and leave when their iteration count is reached."""

    def __init__(self, hbm_budget_bytes: int = 0, values: int = 0, per_pass: int = 4, tile_width: int = 0,
                 tile_rows: int = 0, tile_groups: int = 0, tile_threads: int = 0, use_graph: int = 0):
        self.h = vp()
        o = BatchOpts(values, per_pass, tile_width, tile_rows, tile_groups, tile_threads, use_graph)
        _check(hip_lib().mlem_server_create(hbm_budget_bytes, C.byref(o), C.byref(self.h)))
        self.cols = []
        self.rows = []
        self.matrices = []  # keeps the matrices alive while served

    def add_matrix(self, matrix: "ResidentMatrix", slots: int) -> int:
        mid = u32()
        _check(hip_lib().mlem_server_add_matrix(self.h, matrix.h, slots, C.byref(mid)))
        self.matrices.append(matrix)
        self.rows.append(matrix.rows)
        self.cols.append(matrix.cols)
        return mid.value

    def submit(self, matrix_id: int, g: np.ndarray, f0: np.ndarray, iters: int) -> int:
        g = np.ascontiguousarray(g, dtype=np.int32)
        f0 = np.ascontiguousarray(f0, dtype=np.float32)
        if matrix_id >= len(self.cols) or g.size < self.rows[matrix_id] or f0.size != self.cols[matrix_id]:
            raise MlemError(1, "Server.submit: unknown matrix, or sinogram or image of the wrong size")
        rid = u64()
        _check(hip_lib().mlem_server_submit(self.h, matrix_id, _ptr(g), _ptr(f0), iters, C.byref(rid)))
        return rid.value

    def step(self, max_wait_us: int = 0xFFFFFFFF):
        """One scheduling step (start an iteration on every idle busy batch, wait up to max_wait_us for
        one to end); returns (active, queued) requests after it."""
        a, q = u32(), u32()
        _check(hip_lib().mlem_server_step(self.h, max_wait_us, C.byref(a), C.byref(q)))
        return a.value, q.value

    def poll(self, request_id: int, matrix_id: int):
        """(image, latency_s) once the request is done (it is then released), else None."""
        out = np.empty(self.cols[matrix_id], np.float32)
        lat, done = f64(), cint()
        _check(hip_lib().mlem_server_poll(self.h, request_id, _ptr(out), C.byref(lat), C.byref(done)))
        return (out, lat.value) if done.value else None

    def info(self):
        b, u, m, s = u64(), u64(), u32(), u64()
        _check(hip_lib().mlem_server_info(self.h, C.byref(b), C.byref(u), C.byref(m), C.byref(s)))
        return dict(budget=b.value, used=u.value, matrices=m.value, steps=s.value)

    def stats(self):
        """Passes over a matrix started so far, and requests moved to a lower pass to empty one."""
        sw, mv = u64(), u64()
        _check(hip_lib().mlem_server_stats(self.h, C.byref(sw), C.byref(mv)))
        return dict(pass_sweeps=sw.value, slot_moves=mv.value)

    def close(self):
        if self.h:
            hip_lib().mlem_server_destroy(self.h)
            self.h = vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TuneResult(C.Structure):
    _fields_ = [("tile_width", u32), ("tile_rows", u32), ("tile_groups", u32), ("tile_threads", u32),
                ("candidates", u32), ("fp_ms", f64), ("bp_ms", f64)]


class LoadStats(C.Structure):
    _fields_ = [("bytes", u64), ("open_s", f64), ("stream_s", f64), ("total_s", f64)]


class ResidentMatrix:
    """A system matrix (and optionally its transpose) resident in HBM."""

    def __init__(self, A: Csr, AT: Optional[Csr] = None, device: int = 0):
        self.h = vp()
        self.rows, self.cols, self.nnz = A.rows, A.cols, A.nnz
        _check(hip_lib().mlem_matrix_create(
            device, A.rows, A.cols, A.nnz, _ptr(A.rowptr), _ptr(A.col), _ptr(A.val),
            _ptr(AT.rowptr) if AT is not None else None, _ptr(AT.col) if AT is not None else None,
            _ptr(AT.val) if AT is not None else None, C.byref(self.h)))

    @classmethod
    def tiles_only(cls, A: Csr, tile_width: int = 0, tile_rows: int = 0, values: int = 0, per_pass: int = 1,
                   chunk_bytes: int = 0, device: int = 0):
        """A resident as the tiles of A and A^T only (mlem_matrix_create_tiled), for matrices whose CSR copies do
        not fit beside their tiles: per_pass 1 for reconstructions (opts(spmv=0)), 4 or 8 for Batch with that
        many per pass (the matrix's values)."""
        self = cls.__new__(cls)
        self.h = vp()
        self.rows, self.cols, self.nnz = A.rows, A.cols, A.nnz
        _check(hip_lib().mlem_matrix_create_tiled(device, A.rows, A.cols, A.nnz, _ptr(A.rowptr), _ptr(A.col),
                                                  _ptr(A.val), tile_width, tile_rows, values, per_pass,
                                                  chunk_bytes, C.byref(self.h)))
        return self

    @classmethod
    def load_csr4(cls, path: str, device: int = 0, threads: int = 0):
        """A straight from a csr4 file (streamed through pinned buffers, split on the GPU).
        Returns (matrix, stats dict)."""
        self = cls.__new__(cls)
        self.h = vp()
        st = LoadStats()
        _check(hip_lib().mlem_matrix_load_csr4(device, path.encode(), threads, C.byref(self.h), C.byref(st)))
        info = self.info()
        self.rows, self.cols, self.nnz = info["rows"], info["cols"], info["nnz"]
        return self, dict(bytes=st.bytes, open_s=st.open_s, stream_s=st.stream_s, total_s=st.total_s)

    def download(self) -> Csr:
        """A's resident arrays copied back to the host."""
        rp, ci, va = self.device_arrays()[:3]
        rowptr = np.empty(self.rows + 1, np.uint64)
        col = np.empty(self.nnz, np.uint32)
        val = np.empty(self.nnz, np.float32)
        lib = hip_lib()
        _check(lib.mlem_memcpy_d2h(_ptr(rowptr), rp, rowptr.nbytes, None))
        if self.nnz:
            _check(lib.mlem_memcpy_d2h(_ptr(col), ci, col.nbytes, None))
            _check(lib.mlem_memcpy_d2h(_ptr(val), va, val.nbytes, None))
        return Csr(self.rows, self.cols, rowptr, col, val)

    def build_transpose(self, scratch_bytes: int = 0) -> int:
        """A^T on the device (bit-exact with the host transpose) within scratch_bytes of scratch (0: what is
        free).  Returns the column ranges sorted one at a time (0: one sort)."""
        n = u32()
        _check(hip_lib().mlem_matrix_build_transpose_limited(self.h, scratch_bytes, C.byref(n)))
        return n.value

    def set_row_order(self, order=None):
        """Row order of the batched layouts (mlem_matrix_set_row_order): order[i] = place of row i (a LOR) for
        batches created afterwards; None clears it."""
        if order is None:
            _check(hip_lib().mlem_matrix_set_row_order(self.h, None))
            return
        o = np.ascontiguousarray(order, dtype=np.uint32)
        if o.shape != (self.rows,):
            raise ValueError(f"row order of {o.shape[0] if o.ndim == 1 else o.shape} entries for {self.rows} rows")
        _check(hip_lib().mlem_matrix_set_row_order(self.h, o.ctypes.data_as(vp)))

    def autotune(self, per_pass: int = 1, values: int = 0) -> dict:
        """The fastest blocked-kernel tiling for this matrix (mlem_autotune): tile_width, tile_rows, tile_groups
        for opts() / Batch (and tile_threads for Batch), plus the candidates timed and the chosen one's FP and BP
        ms per launch."""
        r = TuneResult()
        _check(hip_lib().mlem_autotune(self.h, per_pass, values, C.byref(r)))
        return dict(tile_width=r.tile_width, tile_rows=r.tile_rows, tile_groups=r.tile_groups,
                    tile_threads=r.tile_threads, candidates=r.candidates, fp_ms=r.fp_ms, bp_ms=r.bp_ms)

    def info(self):
        r, c, n, t, b = u32(), u32(), u64(), cint(), u64()
        _check(hip_lib().mlem_matrix_info(self.h, C.byref(r), C.byref(c), C.byref(n), C.byref(t), C.byref(b)))
        return dict(rows=r.value, cols=c.value, nnz=n.value, has_transpose=bool(t.value), device_bytes=b.value)

    def device_arrays(self):
        ptrs = [vp() for _ in range(6)]
        _check(hip_lib().mlem_matrix_device_arrays(self.h, *[C.byref(p) for p in ptrs]))
        return ptrs

    def download_transpose(self) -> Csr:
        colptr = np.zeros(self.cols + 1, np.uint64)
        rowT = np.empty(self.nnz, np.uint32)
        valT = np.empty(self.nnz, np.float32)
        _check(hip_lib().mlem_matrix_download_transpose(self.h, _ptr(colptr), _ptr(rowT), _ptr(valT)))
        return Csr(self.cols, self.rows, colptr, rowT, valT)

    def sensitivity(self):
        s32 = np.zeros(self.cols, np.float32)
        s64 = np.zeros(self.cols, np.float64)
        _check(hip_lib().mlem_matrix_sensitivity(self.h, _ptr(s32), _ptr(s64)))
        return s32, s64

    def close(self):
        """Frees the matrix; MlemError (and the matrix kept) while a Recon or Batch of it is open."""
        if self.h:
            _check(hip_lib().mlem_matrix_destroy(self.h))
            self.h = vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Recon:
    """One reconstruction over a resident matrix (its own stream and buffers)."""

    def __init__(self, m: ResidentMatrix, g: np.ndarray, f0: np.ndarray, sens: Optional[np.ndarray] = None,
                 options: Optional[Opts] = None):
        self.m = m
        self.h = vp()
        g = np.ascontiguousarray(g, np.int32)
        f0 = np.ascontiguousarray(f0, np.float32)
        sens = None if sens is None else np.ascontiguousarray(sens, np.float32)
        if g.size != m.rows or f0.size != m.cols or (sens is not None and sens.size != m.cols):
            raise ValueError("counts / image / sensitivity sizes do not match the matrix")
        _check(hip_lib().mlem_recon_create(m.h, _ptr(g), _ptr(sens), _ptr(f0),
                                           C.byref(options) if options is not None else None, C.byref(self.h)))

    def iterate(self, n: int):
        _check(hip_lib().mlem_recon_iterate(self.h, n))

    def sync(self):
        _check(hip_lib().mlem_recon_sync(self.h))

    def image(self) -> np.ndarray:
        out = np.empty(self.m.cols, np.float32)
        _check(hip_lib().mlem_recon_get_image(self.h, _ptr(out)))
        return out

    def set_image(self, f: np.ndarray):
        f = np.ascontiguousarray(f, np.float32)
        _check(hip_lib().mlem_recon_set_image(self.h, _ptr(f)))

    def kernel_times(self):
        a, b, n = f64(), f64(), u32()
        _check(hip_lib().mlem_recon_kernel_times(self.h, C.byref(a), C.byref(b), C.byref(n)))
        return a.value, b.value, n.value

    def phase_times(self):
        """Summed ms of (FP SpMV kernel, FP epilogue, BP SpMV kernel, BP epilogue) and the iterations timed."""
        ms, n = (f64 * 4)(), u32()
        _check(hip_lib().mlem_recon_phase_times(self.h, ms, C.byref(n)))
        return tuple(ms), n.value

    def reset_times(self):
        _check(hip_lib().mlem_recon_reset_times(self.h))

    def loglik(self) -> float:
        v = f64()
        _check(hip_lib().mlem_recon_loglik(self.h, C.byref(v)))
        return v.value

    @property
    def stream(self):
        return hip_lib().mlem_recon_stream(self.h)

    def close(self):
        if self.h:
            hip_lib().mlem_recon_destroy(self.h)
            self.h = vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mlem_run(m: ResidentMatrix, g: np.ndarray, f0: np.ndarray, iters: int, sens: Optional[np.ndarray] = None,
             options: Optional[Opts] = None) -> np.ndarray:
    g = np.ascontiguousarray(g, np.int32)
    f0 = np.ascontiguousarray(f0, np.float32)
    sens = None if sens is None else np.ascontiguousarray(sens, np.float32)
    out = np.empty(m.cols, np.float32)
    _check(hip_lib().mlem_run(m.h, _ptr(g), _ptr(sens), _ptr(f0), _ptr(out), iters,
                              C.byref(options) if options is not None else None))
    return out
