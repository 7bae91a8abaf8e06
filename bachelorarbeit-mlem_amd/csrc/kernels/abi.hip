// C ABI of libmlem_hip.so (include/mlem_hip.h): argument checks, HIP error
// mapping, the resident matrix / reconstruction handles and mlem_run().
//
// The iteration is the loop of main.cu:329-449 on one GPU with the matrix and
// its transpose resident: FP fused with the ratio (calcFwProj_* + calcCorrel),
// BP fused with the update (calcBwProj_* + calcUpdate) -- the reference's
// NCCL allreduce bp->f (main.cu:426-433) is a copy on one GPU and disappears,
// and so do the two memsets (main.cu:436-448) because the kernels overwrite.
// With using_trans == 1 BP is the atomic scatter over A followed by the
// update, which also re-zeroes bp for the next iteration.
#include "abi_internal.hpp"

using namespace mlem;
using namespace mlem::abi;

namespace mlem {
namespace abi {

thread_local std::string g_err;

int set_err(int st, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return st;
}

int hip_err(hipError_t e, const char *what, const char *file, int line) {
    int st = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? MLEM_ERR_OUT_OF_MEMORY
             : (e == hipErrorNoDevice || e == hipErrorInvalidDevice) ? MLEM_ERR_NO_DEVICE
                                                                       : MLEM_ERR_HIP;
    return set_err(st, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
}

// Host-side checks of a CSR's row pointers before it reaches the GPU (a bad
// row pointer would make the kernels read outside their arrays); its column
// indices are checked on the device once uploaded (check_columns_device).
int check_csr_host(const char *what, uint32_t rows, uint32_t cols, uint64_t nnz, const uint64_t *rowptr,
                   const uint32_t *col) {
    (void)cols;
    (void)col;
    if (rowptr[0] != 0 || rowptr[rows] != nnz)
        return set_err(MLEM_ERR_INVALID_ARGUMENT, "%s: row pointers must run from 0 to nnz=%llu", what,
                       (unsigned long long)nnz);
    for (uint32_t r = 0; r < rows; ++r)
        if (rowptr[r + 1] < rowptr[r])
            return set_err(MLEM_ERR_INVALID_ARGUMENT, "%s: row pointers decrease at row %u", what, r);
    return MLEM_OK;
}

int check_columns_device(const char *what, const uint32_t *col, uint64_t nnz, uint32_t cols) {
    if (nnz == 0) return MLEM_OK;
    DevBuf first;
    HIPCHK(first.alloc(sizeof(uint64_t)));
    const uint64_t none = UINT64_MAX;
    HIPCHK(hipMemcpy(first.p, &none, sizeof(none), hipMemcpyHostToDevice));
    HIPCHK(launch_first_bad_column(col, nnz, cols, first.as<uint64_t>(), nullptr));
    uint64_t k = none;
    HIPCHK(hipMemcpy(&k, first.p, sizeof(k), hipMemcpyDeviceToHost));
    if (k == none) return MLEM_OK;
    uint32_t c = 0;
    HIPCHK(hipMemcpy(&c, col + k, sizeof(c), hipMemcpyDeviceToHost));
    return set_err(MLEM_ERR_INVALID_ARGUMENT, "%s: column index %u at nonzero %llu >= %u", what, c,
                   (unsigned long long)k, cols);
}

// Host array -> device: small copies directly, large ones through four
// pinned 64 MiB buffers in rotation, filled by the host library's OpenMP
// threads while the previous buffers go to the device (pageable copies run
// on the calling thread, which OpenMP may have bound to one core).
int upload_h2d(void *dst, const void *src, size_t bytes) {
    constexpr size_t kChunk = size_t(64) << 20;
    if (bytes <= 2 * kChunk) {
        if (bytes) HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
        return MLEM_OK;
    }
    constexpr int kBufs = 4;
    struct Staging {
        void *host[kBufs] = {};
        hipEvent_t done[kBufs] = {};
        hipStream_t s = nullptr;
        ~Staging() {
            if (s) (void)hipStreamSynchronize(s);
            for (int i = 0; i < kBufs; ++i) {
                if (host[i]) (void)hipHostFree(host[i]);
                if (done[i]) (void)hipEventDestroy(done[i]);
            }
            if (s) (void)hipStreamDestroy(s);
        }
    } st;
    HIPCHK(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
    for (int i = 0; i < kBufs; ++i) {
        HIPCHK(hipHostMalloc(&st.host[i], kChunk, hipHostMallocDefault));
        HIPCHK(hipEventCreateWithFlags(&st.done[i], hipEventDisableTiming));
    }
    bool used[kBufs] = {};
    int i = 0;
    for (size_t off = 0; off < bytes; off += kChunk, i = (i + 1) % kBufs) {
        const size_t len = std::min(kChunk, bytes - off);
        if (used[i]) HIPCHK(hipEventSynchronize(st.done[i]));
        if (int e = mlem_host_copy(st.host[i], static_cast<const uint8_t *>(src) + off, len, 0))
            return set_err(e, "upload: %s", mlem_host_last_error());
        HIPCHK(hipMemcpyAsync(static_cast<uint8_t *>(dst) + off, st.host[i], len, hipMemcpyHostToDevice, st.s));
        HIPCHK(hipEventRecord(st.done[i], st.s));
        used[i] = true;
    }
    HIPCHK(hipStreamSynchronize(st.s));
    return MLEM_OK;
}

}  // namespace abi
}  // namespace mlem

// The message mlem_last_error() returns, for the other translation units of
// the library (server.hip).
void mlem::set_last_error(const char *msg) { g_err = msg; }

namespace mlem {
namespace abi {

int get_partition(mlem_matrix *m, int which, uint32_t secsize, const Partition **out) {
    std::lock_guard<std::mutex> lk(m->mu);
    auto key = std::make_pair(which, secsize);
    auto it = m->parts.find(key);
    if (it != m->parts.end()) {
        *out = it->second.get();
        return MLEM_OK;
    }
    auto p = std::make_unique<Partition>();
    const uint32_t nrow = which == 0 ? m->rows : m->cols;
    p->plan = merge_plan(nrow, m->nnz, secsize);
    HIPCHK(p->buf.alloc(merge_partition_bytes(p->plan)));
    const uint64_t *rp = which == 0 ? m->rowptr.as<uint64_t>() : m->rowptrT.as<uint64_t>();
    HIPCHK(launch_merge_partition(p->plan, rp, p->buf.p, nullptr));
    HIPCHK(hipStreamSynchronize(nullptr));
    *out = p.get();
    m->parts.emplace(key, std::move(p));
    return MLEM_OK;
}

// Tiles of A (which 0) or A^T (which 1) with segment width W and R rows per
// block, built on first use.  MLEM_ERR_UNSUPPORTED (and *flags) when the
// matrix does not allow them (a row not sorted by column).  Quad tiles (the
// batched layouts) of a matrix with a row order are built from A with its
// rows in that order, and from that matrix's transpose for A^T.
int get_tiled(mlem_matrix *m, int which, uint32_t W, uint32_t R, uint32_t vt, bool quad, const Tiled **out,
              uint32_t *flags) {
    std::lock_guard<std::mutex> lk(m->mu);
    const bool ordered = quad && m->has_order;
    auto key = std::make_tuple(which, W, R, vt, quad, ordered);
    *flags = 0;
    auto it = m->tiles.find(key);
    if (it != m->tiles.end()) {
        *out = it->second.get();
        return MLEM_OK;
    }
    if (m->tiles_only)
        return set_err(MLEM_ERR_UNSUPPORTED, "a tiles-only matrix holds tiles of width %u and %u rows (values %u) only",
                       m->tile_W, m->tile_R, m->tile_vt);
    auto ref = m->tile_refused.find(key);
    if (ref != m->tile_refused.end()) {
        *flags = ref->second;
        return set_err(MLEM_ERR_UNSUPPORTED, "tiling refused (flags %u)", *flags);
    }
    auto t = std::make_unique<Tiled>();
    t->ordered = ordered;
    const uint32_t nrow = which == 0 ? m->rows : m->cols, ncol = which == 0 ? m->cols : m->rows;
    t->plan = tile_plan(nrow, ncol, m->nnz, W, R, vt, quad);
    const uint64_t *rp = which == 0 ? m->rowptr.as<uint64_t>() : m->rowptrT.as<uint64_t>();
    const uint32_t *ci = which == 0 ? m->col.as<uint32_t>() : m->colT.as<uint32_t>();
    const float *va = which == 0 ? m->val.as<float>() : m->valT.as<float>();
    size_t free_b = 0, total_b = 0;
    const size_t margin = (size_t)1 << 30;
    uint32_t nr = 0;
    // A in the row order (A'), and for A^T its transpose: A^T's columns are
    // then the places of A's rows.  Freed once the tiles are filled.
    DevBuf o_rowptr, o_col, o_val, tcp, trow, o_valT;
    const uint64_t *arp = m->rowptr.as<uint64_t>();  // the A whose codes A^T's 16-bit tiles take
    const uint32_t *aci = m->col.as<uint32_t>();
    const float *ava = m->val.as<float>();
    if (ordered) {
        HIPCHK(o_rowptr.alloc(sizeof(uint64_t) * ((size_t)m->rows + 1)));
        HIPCHK(o_col.alloc(sizeof(uint32_t) * std::max<uint64_t>(1, m->nnz)));
        HIPCHK(o_val.alloc(sizeof(float) * std::max<uint64_t>(1, m->nnz)));
        HIPCHK(reorder_rows(m->rows, m->nnz, m->rowptr.as<uint64_t>(), m->col.as<uint32_t>(), m->val.as<float>(),
                            m->row_order.as<uint32_t>(), o_rowptr.as<uint64_t>(), o_col.as<uint32_t>(),
                            o_val.as<float>(), nullptr));
        arp = o_rowptr.as<uint64_t>();
        aci = o_col.as<uint32_t>();
        ava = o_val.as<float>();
        if (which == 0) {
            rp = arp;
            ci = aci;
            va = ava;
        } else {
            HIPCHK(tcp.alloc(sizeof(uint64_t) * ((size_t)m->cols + 1)));
            HIPCHK(trow.alloc(sizeof(uint32_t) * std::max<uint64_t>(1, m->nnz)));
            rp = tcp.as<uint64_t>();
            ci = trow.as<uint32_t>();
            if (vt == VAL_F32) {
                HIPCHK(o_valT.alloc(sizeof(float) * std::max<uint64_t>(1, m->nnz)));
                HIPCHK(hipMemGetInfo(&free_b, &total_b));
                HIPCHK(transpose_in_ranges(m->rows, m->cols, m->nnz, arp, aci, ava, tcp.as<uint64_t>(),
                                           trow.as<uint32_t>(), o_valT.as<float>(),
                                           free_b > margin ? free_b - margin : 0, &nr, nullptr));
                va = o_valT.as<float>();
            }
        }
    }
    // 16-bit values on A^T: the codes of A's quantization, transposed, stored
    // as they are, so that the forward and back projections use one quantized
    // matrix (and the sensitivity is its column sums).
    DevBuf codesT;
    if (which == 1 && vt != VAL_F32) {
        DevBuf scl, codes, q;
        HIPCHK(scl.alloc(sizeof(float) * std::max<size_t>(1, m->rows)));
        HIPCHK(codes.alloc(sizeof(float) * std::max<uint64_t>(1, m->nnz)));
        HIPCHK(q.alloc(sizeof(uint16_t) * std::max<uint64_t>(1, m->nnz)));
        HIPCHK(launch_quantize_codes(arp, ava, m->rows, m->nnz, vt, scl.as<float>(), codes.as<float>(),
                                     q.as<uint16_t>(), nullptr));
        q.reset();
        o_val.reset();
        DevBuf tcp_l, trow_l;
        if (!ordered) {  // the resident transpose holds the structure; these are scratch
            HIPCHK(tcp_l.alloc(sizeof(uint64_t) * ((size_t)m->cols + 1)));
            HIPCHK(trow_l.alloc(sizeof(uint32_t) * std::max<uint64_t>(1, m->nnz)));
        }
        uint64_t *tc = ordered ? tcp.as<uint64_t>() : tcp_l.as<uint64_t>();
        uint32_t *tr = ordered ? trow.as<uint32_t>() : trow_l.as<uint32_t>();
        HIPCHK(codesT.alloc(sizeof(float) * std::max<uint64_t>(1, m->nnz)));
        HIPCHK(hipMemGetInfo(&free_b, &total_b));
        HIPCHK(transpose_in_ranges(m->rows, m->cols, m->nnz, arp, aci, codes.as<float>(), tc, tr,
                                   codesT.as<float>(), free_b > margin ? free_b - margin : 0, &nr, nullptr));
        HIPCHK(t->sens.alloc(sizeof(float) * std::max<size_t>(1, m->cols)));
        HIPCHK(launch_colsum_weighted_f64(m->rowptrT.as<uint64_t>(), tr, codesT.as<float>(), scl.as<float>(),
                                          m->cols, nullptr, t->sens.as<float>(), nullptr));
        HIPCHK(hipStreamSynchronize(nullptr));
        va = codesT.as<float>();
        t->plan.unit_scale = true;
    }
    DevBuf scratch;
    HIPCHK(scratch.alloc(tile_count_scratch_bytes(t->plan)));
    uint32_t fl = 0;
    HIPCHK(launch_tile_count(t->plan, rp, ci, scratch.p, &fl, nullptr));
    if (fl) {
        m->tile_refused[key] = fl;
        *flags = fl;
        return set_err(MLEM_ERR_UNSUPPORTED, "tiling (width %u, rows %u) refused: a row is not sorted by column", W,
                       R);
    }
    HIPCHK(t->mem.alloc(tiled_matrix_bytes(t->plan)));
    t->view = tiled_matrix_view(t->plan, t->mem.p);
    {
        DevBuf fill_scratch;
        HIPCHK(fill_scratch.alloc(tile_fill_scratch_bytes(t->plan)));
        HIPCHK(launch_tile_fill(t->plan, rp, ci, va, t->mem.p, scratch.p, fill_scratch.p, &fl, nullptr));
    }
    if (fl & 8u) return set_err(MLEM_ERR_INVALID_ARGUMENT, "uf16 values need a matrix without negative values");
    if (fl) return set_err(MLEM_ERR_FORMAT, "tiling (width %u, rows %u): layout check failed (flags %u)", W, R, fl);
    progress("tiles of %s: %u columns x %u rows, %s%s, values %u: %llu entries stored for %llu nonzeros (+%.1f%%)",
             which ? "A^T" : "A", W, R, quad ? "quad groups" : "row steps", ordered ? " in the row order" : "", vt,
             (unsigned long long)(t->plan.cap - 8), (unsigned long long)m->nnz,
             m->nnz ? 100.0 * ((double)(t->plan.cap - 8) / (double)m->nnz - 1.0) : 0.0);
    *out = t.get();
    m->tiles.emplace(key, std::move(t));
    return MLEM_OK;
}

// Progress of long builds on stderr when MLEM_VERBOSE is set (not empty, not 0).
void progress(const char *fmt, ...) {
    static const bool on = [] {
        const char *v = std::getenv("MLEM_VERBOSE");
        return v && *v && std::strcmp(v, "0") != 0;
    }();
    if (!on) return;
    va_list ap;
    va_start(ap, fmt);
    std::fprintf(stderr, "[mlem] ");
    std::vfprintf(stderr, fmt, ap);
    std::fprintf(stderr, "\n");
    std::fflush(stderr);
    va_end(ap);
}

// Work list of tiled matrix t for about `target` workgroups, built on first
// use from the tile sizes.
int get_tile_work(mlem_matrix *m, Tiled *t, uint32_t target, const TileWorkDev **out) {
    std::lock_guard<std::mutex> lk(m->mu);
    const uint32_t key = target;
    auto it = t->works.find(key);
    if (it != t->works.end()) {
        *out = it->second.get();
        return MLEM_OK;
    }
    const TilePlan &p = t->plan;
    std::vector<uint32_t> nz(p.ntiles);
    std::vector<uint64_t> off(p.ntiles + 1);
    if (p.ntiles) {
        HIPCHK(hipMemcpy(nz.data(), t->view.tile_nz, 4 * p.ntiles, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(off.data(), t->view.tile_off, 8 * (p.ntiles + 1), hipMemcpyDeviceToHost));
    }
    const TileWorkList L = tile_work_list(p, nz.data(), off.data(), target);
    if (L.bad) return set_err(MLEM_ERR_FORMAT, "tiles (width %u, rows %u) are not aligned for the stream kernel", p.W, p.R);
    auto w = std::make_unique<TileWorkDev>();
    const size_t b_wl = 4 * L.wl.size(), b_tl = 4 * L.tiles.size(), b_first = 4 * L.wl_first.size(),
                 b_order = 4 * L.order.size();
    const size_t o_tl = (b_wl + 255) & ~size_t(255), o_first = o_tl + ((b_tl + 255) & ~size_t(255)),
                 o_order = o_first + ((b_first + 255) & ~size_t(255));
    HIPCHK(w->mem.alloc(std::max<size_t>(256, o_order + b_order)));
    char *base = static_cast<char *>(w->mem.p);
    if (b_wl) HIPCHK(hipMemcpy(base, L.wl.data(), b_wl, hipMemcpyHostToDevice));
    if (b_tl) HIPCHK(hipMemcpy(base + o_tl, L.tiles.data(), b_tl, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(base + o_first, L.wl_first.data(), b_first, hipMemcpyHostToDevice));
    if (b_order) HIPCHK(hipMemcpy(base + o_order, L.order.data(), b_order, hipMemcpyHostToDevice));
    w->view.span = L.span;
    w->view.nwg = L.nwg;
    w->view.ntl = L.ntl;
    for (size_t b = 0; b + 1 < L.wl_first.size(); ++b)
        w->view.max_runs = std::max(w->view.max_runs, L.wl_first[b + 1] - L.wl_first[b]);
    w->view.wl = reinterpret_cast<const uint4 *>(base);
    w->view.tiles = reinterpret_cast<const uint4 *>(base + o_tl);
    w->view.wl_first = reinterpret_cast<const uint32_t *>(base + o_first);
    w->view.order = b_order ? reinterpret_cast<const uint32_t *>(base + o_order) : nullptr;
    *out = w.get();
    t->works.emplace(key, std::move(w));
    return MLEM_OK;
}

// The blocked path for one side (0: A, 1: A^T): tiles (built on first use),
// a work list of about `groups` runs (0: 24 per CU) and the partial-sum
// buffer for BC right-hand sides.  MLEM_ERR_UNSUPPORTED when the matrix or
// the LDS does not allow the shape.
int setup_tiled(mlem_matrix *m, int which, uint32_t W, uint32_t R, uint32_t vt, uint32_t BC, uint32_t groups,
                const Tiled **tiled, const TileWorkDev **work, DevBuf *part) {
    *tiled = nullptr;
    *work = nullptr;
    if (!valid_tile_config(W, BC))
        return set_err(MLEM_ERR_UNSUPPORTED, "no blocked kernel for width %u with %u right-hand sides", W, BC);
    uint32_t flags = 0;
    const Tiled *t = nullptr;
    if (int st = get_tiled(m, which, W, R, vt, BC > 1, &t, &flags)) return st;
    int dev = 0, lds_max = 65536, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        (void)hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    }
    const uint32_t target = groups ? groups : 24u * (uint32_t)std::max(1, cus);
    const TileWorkDev *w = nullptr;
    if (int st = get_tile_work(m, const_cast<Tiled *>(t), target, &w)) return st;
    if (tiled_lds_bytes(t->plan, w->view.span, BC) + 4096 > (size_t)lds_max)
        return set_err(MLEM_ERR_UNSUPPORTED, "tiling (width %u, rows %u, %u right-hand sides) needs too much LDS", W,
                       R, BC);
    HIPCHK(part->alloc(std::max<size_t>(4, tiled_part_bytes(t->plan, w->view.nwg, BC))));
    *tiled = t;
    *work = w;
    return MLEM_OK;
}

uint32_t default_tile_rows(uint64_t nnz, uint32_t groups, uint32_t BC) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const uint64_t target = groups ? groups : (BC == 1 ? 24u : 8u) * (uint64_t)std::max(1, cus);
    if (nnz < target * 2 * kTileUnit) return 512u;
    return BC == 1 ? 2048u : 1024u;
}

// Chooses the blocked (LDS-tiled) or the plain merge path for one side of
// the iteration.  The blocked path (default tiles of 2048 rows, 512 for
// small matrices, x 1024 columns) runs about tile_groups workgroups (0: 24
// per CU); secsize applies to the plain merge path only.  16-bit values need
// the blocked path.
int setup_merge_side(mlem_matrix *m, int which, const mlem_opts &o, uint32_t secsize, const Tiled **tiled,
                     const TileWorkDev **work, DevBuf *part, const Partition **partn, DevBuf *scratch) {
    *tiled = nullptr;
    *work = nullptr;
    if (o.tile_width != 1) {
        const uint32_t W = o.tile_width == 0 ? 1024u : o.tile_width;
        const uint32_t R = o.tile_rows == 0 ? default_tile_rows(m->nnz, o.tile_groups, 1) : o.tile_rows;
        const int st = setup_tiled(m, which, W, R, o.values, 1, o.tile_groups, tiled, work, part);
        if (st == MLEM_OK || m->tiles_only) return st;
        // Automatic width with f32 values: a matrix the tiles do not suit, or
        // whose tiles do not fit beside the resident copies (the 100 GB
        // matrices), takes the plain merge path; otherwise no fallback.
        if ((st != MLEM_ERR_UNSUPPORTED && st != MLEM_ERR_OUT_OF_MEMORY) || o.tile_width != 0 || o.values != VAL_F32)
            return st;
        (void)hipGetLastError();  // the failed allocation's error is handled here
    }
    if (o.values != VAL_F32)
        return set_err(MLEM_ERR_UNSUPPORTED, "16-bit matrix values need the blocked merge-based kernels");
    if (int st = get_partition(m, which, secsize, partn)) return st;
    HIPCHK(scratch->alloc(merge_scratch_bytes((*partn)->plan)));
    return MLEM_OK;
}

// One MLEM iteration on r->stream (main.cu:331-448).
hipError_t enqueue_iteration(mlem_recon *r, IterEvents *ev) {
    const mlem_matrix *m = r->m;
    hipStream_t s = r->stream;
    hipError_t e;
    auto mark = [&](int i) { return ev ? hipEventRecord(ev->e[i], s) : hipSuccess; };
    if ((e = mark(0)) != hipSuccess) return e;
    // FP + ratio
    EpiArgs fa{};
    fa.out = r->ratio.as<float>();
    fa.g = r->g.as<int32_t>();
    if (r->tfp) {
        if ((e = launch_tiled_partials(r->tfp->plan, r->tfp->view, r->wfp->view, r->f.as<float>(),
                                       r->part_fp.as<float>(), 1, 0, s)) != hipSuccess)
            return e;
        if ((e = mark(1)) != hipSuccess) return e;
        e = launch_tiled_reduce(r->tfp->plan, r->tfp->view, r->wfp->view, r->part_fp.as<float>(), 1, EPI_RATIO, fa, s);
    } else {
        if (r->opts.spmv == 0)
            e = launch_merge_spmv(r->pfp->plan, r->pfp->buf.p, r->scratch_fp.p, m->rowptr.as<uint64_t>(),
                                  m->col.as<uint32_t>(), m->val.as<float>(), r->f.as<float>(), EPI_RATIO, fa, s);
        else
            e = launch_csr_vector(r->lanes_fp, m->rowptr.as<uint64_t>(), m->col.as<uint32_t>(), m->val.as<float>(),
                                  r->f.as<float>(), m->rows, EPI_RATIO, fa, s);
        if (e == hipSuccess) e = mark(1);
    }
    if (e != hipSuccess) return e;
    if ((e = mark(2)) != hipSuccess) return e;
    // BP + update
    EpiArgs ba{};
    ba.norm = r->sens.as<float>();
    ba.f = r->f.as<float>();
    if (r->opts.using_trans == 0) {
        if (r->tbp) {
            if ((e = launch_tiled_partials(r->tbp->plan, r->tbp->view, r->wbp->view, r->ratio.as<float>(),
                                           r->part_bp.as<float>(), 1, 0, s)) != hipSuccess)
                return e;
            if ((e = mark(3)) != hipSuccess) return e;
            e = launch_tiled_reduce(r->tbp->plan, r->tbp->view, r->wbp->view, r->part_bp.as<float>(), 1, EPI_UPDATE,
                                    ba, s);
        } else {
            if (r->bp_merge)
                e = launch_merge_spmv(r->pbp->plan, r->pbp->buf.p, r->scratch_bp.p, m->rowptrT.as<uint64_t>(),
                                      m->colT.as<uint32_t>(), m->valT.as<float>(), r->ratio.as<float>(), EPI_UPDATE,
                                      ba, s);
            else
                e = launch_csr_vector(r->lanes_bp, m->rowptrT.as<uint64_t>(), m->colT.as<uint32_t>(),
                                      m->valT.as<float>(), r->ratio.as<float>(), m->cols, EPI_UPDATE, ba, s);
            if (e == hipSuccess) e = mark(3);
        }
    } else {
        e = launch_scatter(r->lanes_fp, m->rowptr.as<uint64_t>(), m->col.as<uint32_t>(), m->val.as<float>(),
                           r->ratio.as<float>(), r->bp.as<float>(), m->rows, s);
        if (e == hipSuccess) e = mark(3);
        if (e == hipSuccess)
            e = launch_update_image(r->f.as<float>(), r->sens.as<float>(), r->bp.as<float>(), m->cols, s);
    }
    if (e != hipSuccess) return e;
    return mark(4);
}

int collect_times(mlem_recon *r) {
    for (size_t i = 0; i < r->events_used; ++i) {
        HIPCHK(hipEventSynchronize(r->events[i].e[4]));
        for (int k = 0; k < 4; ++k) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, r->events[i].e[k], r->events[i].e[k + 1]));
            r->phase_ms[k] += ms;
        }
        r->timed++;
    }
    r->events_used = 0;
    return MLEM_OK;
}

int destroy_graph(mlem_recon *r) {
    if (r->graph_exec) (void)hipGraphExecDestroy(r->graph_exec);
    if (r->graph) (void)hipGraphDestroy(r->graph);
    r->graph_exec = nullptr;
    r->graph = nullptr;
    return MLEM_OK;
}

int sens_on_device(const mlem_matrix *m, float *s32, double *s64_or_null, hipStream_t s) {
    DevBuf tmp;
    double *s64 = s64_or_null;
    if (m->tiles_only) {
        if (s32) HIPCHK(hipMemcpyAsync(s32, m->sens32.p, m->sens32.bytes, hipMemcpyDeviceToDevice, s));
        if (s64) HIPCHK(hipMemcpyAsync(s64, m->sens64.p, m->sens64.bytes, hipMemcpyDeviceToDevice, s));
    } else if (m->has_T) {
        HIPCHK(launch_colsum_rows_f64(m->rowptrT.as<uint64_t>(), m->valT.as<float>(), m->cols, s64, s32, s));
    } else {
        if (!s64) {
            HIPCHK(tmp.alloc(sizeof(double) * std::max<size_t>(1, m->cols)));
            s64 = tmp.as<double>();
        }
        HIPCHK(launch_fill_f64(s64, 0.0, m->cols, s));
        HIPCHK(launch_colsum_scatter_f64(m->rowptr.as<uint64_t>(), m->col.as<uint32_t>(), m->val.as<float>(), m->rows,
                                         s64, s));
        if (s32) HIPCHK(launch_f64_to_f32(s64, s32, m->cols, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    return MLEM_OK;
}

}  // namespace abi
}  // namespace mlem

extern "C" {

const char *mlem_last_error(void) { return g_err.c_str(); }

const char *mlem_status_string(int status) {
    switch (status) {
        case MLEM_OK: return "ok";
        case MLEM_ERR_INVALID_ARGUMENT: return "invalid argument";
        case MLEM_ERR_HIP: return "HIP runtime error";
        case MLEM_ERR_OUT_OF_MEMORY: return "out of memory";
        case MLEM_ERR_UNSUPPORTED: return "unsupported";
        case MLEM_ERR_NO_DEVICE: return "no device";
        case MLEM_ERR_IO: return "I/O error";
        case MLEM_ERR_FORMAT: return "format error";
        default: return "unknown status";
    }
}

const char *mlem_version(void) { return "mlem-mi355x 0.1 (gfx950)"; }

// ---------------------------------------------------------------- runtime
int mlem_device_count(int *count) {
    ARGCHK(count, "mlem_device_count: null count");
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e == hipErrorNoDevice) {
        *count = 0;
        return MLEM_OK;
    }
    HIPCHK(e);
    return MLEM_OK;
}

int mlem_set_device(int device) {
    HIPCHK(hipSetDevice(device));
    return MLEM_OK;
}

int mlem_device_info(int device, char *name, size_t name_len, uint64_t *total_mem, int *cu_count) {
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, device));
    if (name && name_len) {  // the marketing name can be empty (no device-id table): the architecture alone
        if (p.name[0])
            std::snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
        else
            std::snprintf(name, name_len, "%s", p.gcnArchName);
    }
    if (total_mem) *total_mem = p.totalGlobalMem;
    if (cu_count) *cu_count = p.multiProcessorCount;
    return MLEM_OK;
}

int mlem_device_alloc(void **ptr, size_t bytes) {
    ARGCHK(ptr, "mlem_device_alloc: null ptr");
    *ptr = nullptr;
    if (bytes == 0) return MLEM_OK;
    HIPCHK(hipMalloc(ptr, bytes));
    return MLEM_OK;
}

int mlem_device_free(void *ptr) {
    if (ptr) HIPCHK(hipFree(ptr));
    return MLEM_OK;
}

int mlem_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return MLEM_OK;
    ARGCHK(dst && src, "mlem_memcpy_h2d: null pointer");
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(stream)));
    HIPCHK(hipStreamSynchronize(S(stream)));
    return MLEM_OK;
}

int mlem_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return MLEM_OK;
    ARGCHK(dst && src, "mlem_memcpy_d2h: null pointer");
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(stream)));
    HIPCHK(hipStreamSynchronize(S(stream)));
    return MLEM_OK;
}

int mlem_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return MLEM_OK;
    ARGCHK(dst && src, "mlem_memcpy_d2d: null pointer");
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(stream)));
    return MLEM_OK;
}

int mlem_stream_create(void **stream) {
    ARGCHK(stream, "mlem_stream_create: null stream");
    hipStream_t s;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return MLEM_OK;
}

int mlem_stream_destroy(void *stream) {
    if (stream) HIPCHK(hipStreamDestroy(S(stream)));
    return MLEM_OK;
}

int mlem_stream_sync(void *stream) {
    HIPCHK(hipStreamSynchronize(S(stream)));
    return MLEM_OK;
}

int mlem_device_sync(void) {
    HIPCHK(hipDeviceSynchronize());
    return MLEM_OK;
}

int mlem_event_create(void **event) {
    ARGCHK(event, "mlem_event_create: null event");
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    *event = e;
    return MLEM_OK;
}

int mlem_event_destroy(void *event) {
    if (event) HIPCHK(hipEventDestroy(static_cast<hipEvent_t>(event)));
    return MLEM_OK;
}

int mlem_event_record(void *event, void *stream) {
    ARGCHK(event, "mlem_event_record: null event");
    HIPCHK(hipEventRecord(static_cast<hipEvent_t>(event), S(stream)));
    return MLEM_OK;
}

int mlem_event_elapsed_ms(void *start, void *stop, float *ms) {
    ARGCHK(start && stop && ms, "mlem_event_elapsed_ms: null argument");
    HIPCHK(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
    HIPCHK(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
    return MLEM_OK;
}

int mlem_fill_f32(float *dst, float value, uint64_t n, void *stream) {
    ARGCHK(dst || n == 0, "mlem_fill_f32: null dst");
    HIPCHK(launch_fill(dst, value, n, S(stream)));
    return MLEM_OK;
}

int mlem_copy_f32(float *dst, const float *src, uint64_t n, void *stream) {
    ARGCHK(n == 0 || (dst && src), "mlem_copy_f32: null pointer");
    ARGCHK(aligned16(dst) && aligned16(src), "mlem_copy_f32: pointers must be 16-byte aligned");
    HIPCHK(launch_copy(dst, src, n, S(stream)));
    return MLEM_OK;
}

// ---------------------------------------------------------------- kernels
int mlem_ratio(const int32_t *g, float *fwproj, uint32_t rows, void *stream) {
    ARGCHK(rows == 0 || (g && fwproj), "mlem_ratio: null pointer");
    HIPCHK(launch_ratio(g, fwproj, rows, S(stream)));
    return MLEM_OK;
}

int mlem_update(const float *f, const float *norm, float *bwproj, uint32_t cols, void *stream) {
    ARGCHK(cols == 0 || (f && norm && bwproj), "mlem_update: null pointer");
    HIPCHK(launch_update(f, norm, bwproj, cols, S(stream)));
    return MLEM_OK;
}

size_t mlem_merge_workspace_bytes(uint32_t rows, uint64_t nnz, uint32_t secsize) {
    if (secsize < 1 || secsize > kMaxSecsize) return 0;
    MergePlan p = merge_plan(rows, nnz, secsize);
    return merge_partition_bytes(p) + merge_scratch_bytes(p);
}

static int merge_call(const char *what, const uint64_t *rowptr, const uint32_t *col, const float *val,
                      const float *x, uint32_t secsize, uint32_t rows, uint64_t nnz, void *ws, size_t ws_bytes,
                      Epi epi, const EpiArgs &ea, void *stream) {
    ARGCHK(secsize >= 1 && secsize <= kMaxSecsize, "%s: secsize %u outside 1..%u", what, secsize, kMaxSecsize);
    if (rows == 0) return MLEM_OK;
    ARGCHK(rowptr && x && (nnz == 0 || (col && val)), "%s: null pointer", what);
    ARGCHK(aligned16(col) && aligned16(val), "%s: col and val must be 16-byte aligned", what);
    MergePlan p = merge_plan(rows, nnz, secsize);
    const size_t need = merge_partition_bytes(p) + merge_scratch_bytes(p);
    ARGCHK(ws && ws_bytes >= need, "%s: workspace of %zu bytes needed (mlem_merge_workspace_bytes)", what, need);
    HIPCHK(launch_merge_partition(p, rowptr, ws, S(stream)));
    HIPCHK(launch_merge_spmv(p, ws, static_cast<char *>(ws) + merge_partition_bytes(p), rowptr, col, val, x, epi, ea,
                             S(stream)));
    return MLEM_OK;
}

int mlem_fp_merge(const uint64_t *rowptr, const uint32_t *col, const float *val, const float *f, float *fwproj,
                  uint32_t secsize, uint32_t rows, uint64_t nnz, void *workspace, size_t workspace_bytes,
                  void *stream) {
    EpiArgs ea{};
    ea.out = fwproj;
    ARGCHK(rows == 0 || fwproj, "mlem_fp_merge: null output");
    return merge_call("mlem_fp_merge", rowptr, col, val, f, secsize, rows, nnz, workspace, workspace_bytes, EPI_PLAIN,
                      ea, stream);
}

int mlem_bp_merge(const uint64_t *rowptrT, const uint32_t *colT, const float *valT, const float *correl,
                  float *bwproj, uint32_t secsize, uint32_t cols, uint64_t nnz, void *workspace,
                  size_t workspace_bytes, void *stream) {
    EpiArgs ea{};
    ea.out = bwproj;
    ARGCHK(cols == 0 || bwproj, "mlem_bp_merge: null output");
    return merge_call("mlem_bp_merge", rowptrT, colT, valT, correl, secsize, cols, nnz, workspace, workspace_bytes,
                      EPI_PLAIN, ea, stream);
}

int mlem_fp_ratio_merge(const uint64_t *rowptr, const uint32_t *col, const float *val, const float *f,
                        const int32_t *g, float *ratio, uint32_t secsize, uint32_t rows, uint64_t nnz,
                        void *workspace, size_t workspace_bytes, void *stream) {
    EpiArgs ea{};
    ea.out = ratio;
    ea.g = g;
    ARGCHK(rows == 0 || (ratio && g), "mlem_fp_ratio_merge: null pointer");
    return merge_call("mlem_fp_ratio_merge", rowptr, col, val, f, secsize, rows, nnz, workspace, workspace_bytes,
                      EPI_RATIO, ea, stream);
}

int mlem_bp_update_merge(const uint64_t *rowptrT, const uint32_t *colT, const float *valT, const float *ratio,
                         const float *norm, float *f, uint32_t secsize, uint32_t cols, uint64_t nnz,
                         void *workspace, size_t workspace_bytes, void *stream) {
    EpiArgs ea{};
    ea.norm = norm;
    ea.f = f;
    ARGCHK(cols == 0 || (norm && f), "mlem_bp_update_merge: null pointer");
    return merge_call("mlem_bp_update_merge", rowptrT, colT, valT, ratio, secsize, cols, nnz, workspace,
                      workspace_bytes, EPI_UPDATE, ea, stream);
}

static int vector_call(const char *what, const uint64_t *rowptr, const uint32_t *col, const float *val,
                       const float *x, uint32_t rows, uint64_t nnz, uint32_t lanes, Epi epi, const EpiArgs &ea,
                       void *stream) {
    if (rows == 0) return MLEM_OK;
    ARGCHK(rowptr && x && (nnz == 0 || (col && val)), "%s: null pointer", what);
    ARGCHK(aligned16(col) && aligned16(val), "%s: col and val must be 16-byte aligned", what);
    if (lanes == 0) lanes = auto_lanes(rows, nnz);
    ARGCHK(valid_lanes(lanes), "%s: lanes must be 4, 8, 16, 32 or 64 (got %u)", what, lanes);
    HIPCHK(launch_csr_vector(lanes, rowptr, col, val, x, rows, epi, ea, S(stream)));
    return MLEM_OK;
}

int mlem_fp_csr_vector(const uint64_t *rowptr, const uint32_t *col, const float *val, const float *f, float *fwproj,
                       uint32_t rows, uint64_t nnz, uint32_t lanes, void *stream) {
    EpiArgs ea{};
    ea.out = fwproj;
    ARGCHK(rows == 0 || fwproj, "mlem_fp_csr_vector: null output");
    return vector_call("mlem_fp_csr_vector", rowptr, col, val, f, rows, nnz, lanes, EPI_PLAIN, ea, stream);
}

int mlem_bp_csr_vector(const uint64_t *rowptrT, const uint32_t *colT, const float *valT, const float *correl,
                       float *bwproj, uint32_t cols, uint64_t nnz, uint32_t lanes, void *stream) {
    EpiArgs ea{};
    ea.out = bwproj;
    ARGCHK(cols == 0 || bwproj, "mlem_bp_csr_vector: null output");
    return vector_call("mlem_bp_csr_vector", rowptrT, colT, valT, correl, cols, nnz, lanes, EPI_PLAIN, ea, stream);
}

int mlem_fp_ratio_csr_vector(const uint64_t *rowptr, const uint32_t *col, const float *val, const float *f,
                             const int32_t *g, float *ratio, uint32_t rows, uint64_t nnz, uint32_t lanes,
                             void *stream) {
    EpiArgs ea{};
    ea.out = ratio;
    ea.g = g;
    ARGCHK(rows == 0 || (ratio && g), "mlem_fp_ratio_csr_vector: null pointer");
    return vector_call("mlem_fp_ratio_csr_vector", rowptr, col, val, f, rows, nnz, lanes, EPI_RATIO, ea, stream);
}

int mlem_bp_update_csr_vector(const uint64_t *rowptrT, const uint32_t *colT, const float *valT, const float *ratio,
                              const float *norm, float *f, uint32_t cols, uint64_t nnz, uint32_t lanes,
                              void *stream) {
    EpiArgs ea{};
    ea.norm = norm;
    ea.f = f;
    ARGCHK(cols == 0 || (norm && f), "mlem_bp_update_csr_vector: null pointer");
    return vector_call("mlem_bp_update_csr_vector", rowptrT, colT, valT, ratio, cols, nnz, lanes, EPI_UPDATE, ea,
                       stream);
}

int mlem_bp_scatter(const uint64_t *rowptr, const uint32_t *col, const float *val, const float *correl,
                    float *bwproj, uint32_t rows, uint32_t cols, uint64_t nnz, uint32_t lanes, void *stream) {
    ARGCHK(cols == 0 || bwproj, "mlem_bp_scatter: null output");
    HIPCHK(launch_fill(bwproj, 0.0f, cols, S(stream)));
    if (rows == 0) return MLEM_OK;
    ARGCHK(rowptr && correl && (nnz == 0 || (col && val)), "mlem_bp_scatter: null pointer");
    if (lanes == 0) lanes = auto_lanes(rows, nnz);
    ARGCHK(valid_lanes(lanes), "mlem_bp_scatter: lanes must be 4, 8, 16, 32 or 64 (got %u)", lanes);
    HIPCHK(launch_scatter(lanes, rowptr, col, val, correl, bwproj, rows, S(stream)));
    return MLEM_OK;
}

int mlem_colsum_f64(const uint64_t *rowptrT, const float *valT, uint32_t cols, uint64_t nnz, double *sens64,
                    float *sens32, void *stream) {
    if (cols == 0) return MLEM_OK;
    ARGCHK(rowptrT && (nnz == 0 || valT), "mlem_colsum_f64: null pointer");
    HIPCHK(launch_colsum_rows_f64(rowptrT, valT, cols, sens64, sens32, S(stream)));
    return MLEM_OK;
}

int mlem_colsum_scatter_f64(const uint64_t *rowptr, const uint32_t *col, const float *val, uint32_t rows,
                            uint32_t cols, uint64_t nnz, double *sens64, float *sens32, void *stream) {
    if (cols == 0) return MLEM_OK;
    ARGCHK(rowptr && (nnz == 0 || (col && val)), "mlem_colsum_scatter_f64: null pointer");
    DevBuf tmp;
    double *s64 = sens64;
    if (!s64) {
        HIPCHK(tmp.alloc(sizeof(double) * cols));
        s64 = tmp.as<double>();
    }
    HIPCHK(launch_fill_f64(s64, 0.0, cols, S(stream)));
    HIPCHK(launch_colsum_scatter_f64(rowptr, col, val, rows, s64, S(stream)));
    if (sens32) HIPCHK(launch_f64_to_f32(s64, sens32, cols, S(stream)));
    if (!sens64) HIPCHK(hipStreamSynchronize(S(stream)));  // before tmp is freed
    return MLEM_OK;
}

// ---------------------------------------------------------------- matrix
void mlem_opts_default(mlem_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->spmv = 0;
    o->using_trans = 0;
    o->secsize_fp = 8;
    o->secsize_bp = 8;
}

int mlem_matrix_create(int device, uint32_t rows, uint32_t cols, uint64_t nnz, const uint64_t *rowptr,
                       const uint32_t *col, const float *val, const uint64_t *rowptrT, const uint32_t *colT,
                       const float *valT, mlem_matrix **out) {
    ARGCHK(out, "mlem_matrix_create: null out");
    *out = nullptr;
    ARGCHK(rowptr && (nnz == 0 || (col && val)), "mlem_matrix_create: null matrix array");
    ARGCHK(cols > 0, "mlem_matrix_create: cols must be > 0");
    if (int st = check_csr_host("mlem_matrix_create (A)", rows, cols, nnz, rowptr, col)) return st;
    if (rowptrT) {
        ARGCHK(nnz == 0 || (colT && valT), "mlem_matrix_create: transpose arrays incomplete");
        if (int st = check_csr_host("mlem_matrix_create (A^T)", cols, rows, nnz, rowptrT, colT)) return st;
    }
    HIPCHK(hipSetDevice(device));
    auto m = std::make_unique<mlem_matrix>();
    m->device = device;
    m->rows = rows;
    m->cols = cols;
    m->nnz = nnz;
    HIPCHK(m->rowptr.alloc(sizeof(uint64_t) * ((size_t)rows + 1)));
    HIPCHK(m->col.alloc(sizeof(uint32_t) * nnz));
    HIPCHK(m->val.alloc(sizeof(float) * nnz));
    if (int st = upload_h2d(m->rowptr.p, rowptr, m->rowptr.bytes)) return st;
    if (nnz) {
        if (int st = upload_h2d(m->col.p, col, m->col.bytes)) return st;
        if (int st = upload_h2d(m->val.p, val, m->val.bytes)) return st;
    }
    if (int st = check_columns_device("mlem_matrix_create (A)", m->col.as<uint32_t>(), nnz, cols)) return st;
    if (rowptrT) {
        HIPCHK(m->rowptrT.alloc(sizeof(uint64_t) * ((size_t)cols + 1)));
        HIPCHK(m->colT.alloc(sizeof(uint32_t) * nnz));
        HIPCHK(m->valT.alloc(sizeof(float) * nnz));
        if (int st = upload_h2d(m->rowptrT.p, rowptrT, m->rowptrT.bytes)) return st;
        if (nnz) {
            if (int st = upload_h2d(m->colT.p, colT, m->colT.bytes)) return st;
            if (int st = upload_h2d(m->valT.p, valT, m->valT.bytes)) return st;
        }
        if (int st = check_columns_device("mlem_matrix_create (A^T)", m->colT.as<uint32_t>(), nnz, rows)) return st;
        m->has_T = true;
    }
    *out = m.release();
    return MLEM_OK;
}

int mlem_matrix_load_csr4(int device, const char *path, int threads, mlem_matrix **out, mlem_load_stats *stats) {
    ARGCHK(path && out, "mlem_matrix_load_csr4: null argument");
    *out = nullptr;
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto secs = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
    mlem_csr4 *f = nullptr;
    if (int st = mlem_csr4_open(path, &f)) return set_err(st, "%s", mlem_host_last_error());
    struct Closer {
        mlem_csr4 *f;
        ~Closer() { mlem_csr4_close(f); }
    } closer{f};
    uint32_t rows = 0, cols = 0;
    uint64_t nnz = 0;
    mlem_csr4_info(f, &rows, &cols, &nnz, nullptr, nullptr);
    ARGCHK(cols > 0, "mlem_matrix_load_csr4: %s has no columns", path);
    const uint64_t *ends = mlem_csr4_row_ends(f);
    const uint8_t *data = static_cast<const uint8_t *>(mlem_csr4_data(f));
    const auto t1 = clk::now();

    HIPCHK(hipSetDevice(device));
    auto m = std::make_unique<mlem_matrix>();
    m->device = device;
    m->rows = rows;
    m->cols = cols;
    m->nnz = nnz;
    HIPCHK(m->rowptr.alloc(sizeof(uint64_t) * ((size_t)rows + 1)));
    HIPCHK(m->col.alloc(sizeof(uint32_t) * nnz));
    HIPCHK(m->val.alloc(sizeof(float) * nnz));
    {
        std::vector<uint64_t> rp((size_t)rows + 1, 0);
        if (rows) std::memcpy(rp.data() + 1, ends, sizeof(uint64_t) * rows);
        HIPCHK(hipMemcpy(m->rowptr.p, rp.data(), m->rowptr.bytes, hipMemcpyHostToDevice));
    }
    // Four pinned buffers of 128 MiB in rotation: while buffer i goes to the
    // device (copy stream) and is split into col / val (split stream), the
    // host library's OpenMP threads fill buffer i+1 from the mapped file.
    // (Threads started here would inherit the one core that OpenMP binds a
    // process's initial thread to under OMP_PROC_BIND: a fifth of the rate.)
    constexpr int kBufs = 4;
    constexpr size_t kChunk = size_t(128) << 20;  // a multiple of the 8-byte entry
    const uint64_t total = 8 * nnz;
    // Declared before the streams: on an early return the streams are
    // synchronized (Staging's destructor) before these are freed.
    DevBuf dev[kBufs], bad;
    struct Staging {
        void *host[kBufs] = {};
        hipEvent_t copied[kBufs] = {}, split[kBufs] = {};
        hipStream_t s = nullptr, c = nullptr;  // split stream, copy stream
        ~Staging() {
            if (c) (void)hipStreamSynchronize(c);
            if (s) (void)hipStreamSynchronize(s);
            for (int i = 0; i < kBufs; ++i) {
                if (host[i]) (void)hipHostFree(host[i]);
                if (copied[i]) (void)hipEventDestroy(copied[i]);
                if (split[i]) (void)hipEventDestroy(split[i]);
            }
            if (c) (void)hipStreamDestroy(c);
            if (s) (void)hipStreamDestroy(s);
        }
    } st;
    HIPCHK(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&st.c, hipStreamNonBlocking));
    HIPCHK(bad.alloc(sizeof(uint32_t)));
    HIPCHK(launch_fill_u32(bad.as<uint32_t>(), 0u, 1, st.s));
    if (total) {
        for (int i = 0; i < kBufs; ++i) {
            HIPCHK(hipHostMalloc(&st.host[i], kChunk, hipHostMallocDefault));
            HIPCHK(hipEventCreateWithFlags(&st.copied[i], hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&st.split[i], hipEventDisableTiming));
            HIPCHK(dev[i].alloc(kChunk));
        }
    }
    double t_fill = 0, t_wait = 0;
    bool used[kBufs] = {};
    int i = 0;
    for (uint64_t off = 0; off < total; off += kChunk, i = (i + 1) % kBufs) {
        const size_t len = (size_t)std::min<uint64_t>(kChunk, total - off);
        const auto w0 = clk::now();
        if (used[i]) HIPCHK(hipEventSynchronize(st.split[i]));  // the buffer's last chunk is on the device and split
        const auto w1 = clk::now();
        if (int e = mlem_host_copy(st.host[i], data + off, len, threads))
            return set_err(e, "mlem_matrix_load_csr4: %s", mlem_host_last_error());
        const auto w2 = clk::now();
        t_wait += secs(w0, w1);
        t_fill += secs(w1, w2);
        HIPCHK(hipMemcpyAsync(dev[i].p, st.host[i], len, hipMemcpyHostToDevice, st.c));
        HIPCHK(hipEventRecord(st.copied[i], st.c));
        HIPCHK(hipStreamWaitEvent(st.s, st.copied[i], 0));
        HIPCHK(launch_aos_split(dev[i].p, len / 8, off / 8, cols, m->col.as<uint32_t>(), m->val.as<float>(),
                                bad.as<uint32_t>(), st.s));
        HIPCHK(hipEventRecord(st.split[i], st.s));
        used[i] = true;
    }
    uint32_t flag = 0;
    HIPCHK(hipMemcpyAsync(&flag, bad.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st.s));
    HIPCHK(hipStreamSynchronize(st.s));
    const auto t2 = clk::now();
    progress("csr4 load: %.2f GB in %.3f s: host threads filling the pinned buffers %.3f s, waiting for a buffer "
             "%.3f s (%d buffers of %zu MiB)",
             total / 1e9, secs(t1, t2), t_fill, t_wait, kBufs, kChunk >> 20);
    if (flag) return set_err(MLEM_ERR_FORMAT, "mlem_matrix_load_csr4: %s has a column index >= %u", path, cols);
    if (stats) {
        stats->bytes = total;
        stats->open_s = secs(t0, t1);
        stats->stream_s = secs(t1, t2);
        stats->total_s = secs(t0, t2);
    }
    *out = m.release();
    return MLEM_OK;
}

int mlem_matrix_build_transpose(mlem_matrix *m) { return mlem_matrix_build_transpose_limited(m, 0, nullptr); }

int mlem_matrix_build_transpose_limited(mlem_matrix *m, uint64_t scratch_bytes, uint32_t *ranges) {
    ARGCHK(m, "mlem_matrix_build_transpose: null matrix");
    ARGCHK(!m->tiles_only, "mlem_matrix_build_transpose: a tiles-only matrix has no CSR");
    ARGCHK(m->users.load() == 0,
           "mlem_matrix_build_transpose: %u reconstructions or batches still use the matrix's transpose",
           m->users.load());
    HIPCHK(hipSetDevice(m->device));
    if (ranges) *ranges = 0;
    DevBuf scratch, rp, ci, va;
    HIPCHK(rp.alloc(sizeof(uint64_t) * ((size_t)m->cols + 1)));
    HIPCHK(ci.alloc(sizeof(uint32_t) * m->nnz));
    HIPCHK(va.alloc(sizeof(float) * m->nnz));
    const size_t whole = transpose_scratch_bytes(m->rows, m->cols, m->nnz);
    size_t budget = scratch_bytes;
    if (budget == 0) {  // automatic: one sort when its scratch fits beside a margin, else ranges in what is free
        size_t free_b = 0, total_b = 0;
        HIPCHK(hipMemGetInfo(&free_b, &total_b));
        const size_t margin = (size_t)1 << 30;
        budget = free_b > margin ? free_b - margin : 0;
    }
    if (whole <= budget) {
        HIPCHK(scratch.alloc(whole));
        HIPCHK(launch_transpose(m->rows, m->cols, m->nnz, m->rowptr.as<uint64_t>(), m->col.as<uint32_t>(),
                                m->val.as<float>(), rp.as<uint64_t>(), ci.as<uint32_t>(), va.as<float>(), scratch.p,
                                nullptr));
        HIPCHK(hipStreamSynchronize(nullptr));
    } else {
        uint32_t nr = 0;
        const hipError_t e = transpose_in_ranges(m->rows, m->cols, m->nnz, m->rowptr.as<uint64_t>(),
                                                 m->col.as<uint32_t>(), m->val.as<float>(), rp.as<uint64_t>(),
                                                 ci.as<uint32_t>(), va.as<float>(), budget, &nr, nullptr);
        if (e == hipErrorOutOfMemory)
            return set_err(MLEM_ERR_OUT_OF_MEMORY,
                           "mlem_matrix_build_transpose: %zu bytes of scratch cannot hold one column range", budget);
        HIPCHK(e);
        if (ranges) *ranges = nr;
    }
    std::lock_guard<std::mutex> lk(m->mu);
    std::swap(m->rowptrT.p, rp.p);
    std::swap(m->rowptrT.bytes, rp.bytes);
    std::swap(m->colT.p, ci.p);
    std::swap(m->colT.bytes, ci.bytes);
    std::swap(m->valT.p, va.p);
    std::swap(m->valT.bytes, va.bytes);
    m->has_T = true;
    for (auto it = m->parts.begin(); it != m->parts.end();)
        it = it->first.first == 1 ? m->parts.erase(it) : std::next(it);
    for (auto it = m->tiles.begin(); it != m->tiles.end();)
        it = std::get<0>(it->first) == 1 ? m->tiles.erase(it) : std::next(it);
    for (auto it = m->tile_refused.begin(); it != m->tile_refused.end();)
        it = std::get<0>(it->first) == 1 ? m->tile_refused.erase(it) : std::next(it);
    return MLEM_OK;
}

int mlem_matrix_set_row_order(mlem_matrix *m, const uint32_t *order) {
    ARGCHK(m, "mlem_matrix_set_row_order: null matrix");
    ARGCHK(!m->tiles_only, "mlem_matrix_set_row_order: a tiles-only matrix has no CSR to reorder");
    ARGCHK(m->users.load() == 0, "mlem_matrix_set_row_order: %u reconstructions or batches still use the matrix",
           m->users.load());
    HIPCHK(hipSetDevice(m->device));
    DevBuf buf;
    if (order) {
        std::vector<uint8_t> seen(m->rows, 0);
        for (uint32_t i = 0; i < m->rows; ++i) {
            ARGCHK(order[i] < m->rows && !seen[order[i]],
                   "mlem_matrix_set_row_order: not a permutation of the %u rows (entry %u is %u)", m->rows, i,
                   order[i]);
            seen[order[i]] = 1;
        }
        HIPCHK(buf.alloc(sizeof(uint32_t) * std::max<uint32_t>(1, m->rows)));
        if (m->rows) HIPCHK(hipMemcpy(buf.p, order, sizeof(uint32_t) * m->rows, hipMemcpyHostToDevice));
    }
    std::lock_guard<std::mutex> lk(m->mu);
    for (auto it = m->tiles.begin(); it != m->tiles.end();)  // tiles built in the previous order
        it = std::get<5>(it->first) ? m->tiles.erase(it) : std::next(it);
    for (auto it = m->tile_refused.begin(); it != m->tile_refused.end();)
        it = std::get<5>(it->first) ? m->tile_refused.erase(it) : std::next(it);
    std::swap(m->row_order.p, buf.p);
    std::swap(m->row_order.bytes, buf.bytes);
    m->has_order = order != nullptr;
    return MLEM_OK;
}

int mlem_matrix_destroy(mlem_matrix *m) {
    if (!m) return MLEM_OK;
    ARGCHK(m->users.load() == 0, "mlem_matrix_destroy: %u reconstructions or batches still use the matrix",
           m->users.load());
    (void)hipSetDevice(m->device);
    delete m;
    return MLEM_OK;
}

int mlem_matrix_info(const mlem_matrix *m, uint32_t *rows, uint32_t *cols, uint64_t *nnz, int *has_transpose,
                     uint64_t *device_bytes) {
    ARGCHK(m, "mlem_matrix_info: null matrix");
    if (rows) *rows = m->rows;
    if (cols) *cols = m->cols;
    if (nnz) *nnz = m->nnz;
    if (has_transpose) *has_transpose = m->has_T ? 1 : 0;
    if (device_bytes) *device_bytes = m->device_bytes();
    return MLEM_OK;
}

int mlem_matrix_device_arrays(const mlem_matrix *m, const uint64_t **rowptr, const uint32_t **col, const float **val,
                              const uint64_t **rowptrT, const uint32_t **colT, const float **valT) {
    ARGCHK(m, "mlem_matrix_device_arrays: null matrix");
    ARGCHK(!m->tiles_only, "mlem_matrix_device_arrays: a tiles-only matrix has no CSR arrays");
    if (rowptr) *rowptr = m->rowptr.as<uint64_t>();
    if (col) *col = m->col.as<uint32_t>();
    if (val) *val = m->val.as<float>();
    if (rowptrT) *rowptrT = m->rowptrT.as<uint64_t>();
    if (colT) *colT = m->colT.as<uint32_t>();
    if (valT) *valT = m->valT.as<float>();
    return MLEM_OK;
}

int mlem_matrix_download_transpose(const mlem_matrix *m, uint64_t *colptrT, uint32_t *rowT, float *valT) {
    ARGCHK(m, "mlem_matrix_download_transpose: null matrix");
    ARGCHK(m->has_T, "mlem_matrix_download_transpose: matrix has no transpose");
    HIPCHK(hipSetDevice(m->device));
    if (colptrT) HIPCHK(hipMemcpy(colptrT, m->rowptrT.p, m->rowptrT.bytes, hipMemcpyDeviceToHost));
    if (rowT && m->nnz) HIPCHK(hipMemcpy(rowT, m->colT.p, m->colT.bytes, hipMemcpyDeviceToHost));
    if (valT && m->nnz) HIPCHK(hipMemcpy(valT, m->valT.p, m->valT.bytes, hipMemcpyDeviceToHost));
    return MLEM_OK;
}

int mlem_matrix_sensitivity(const mlem_matrix *m, float *sens32, double *sens64) {
    ARGCHK(m, "mlem_matrix_sensitivity: null matrix");
    HIPCHK(hipSetDevice(m->device));
    DevBuf d32, d64;
    HIPCHK(d32.alloc(sizeof(float) * m->cols));
    HIPCHK(d64.alloc(sizeof(double) * m->cols));
    if (int st = sens_on_device(m, d32.as<float>(), d64.as<double>(), nullptr)) return st;
    if (sens32) HIPCHK(hipMemcpy(sens32, d32.p, d32.bytes, hipMemcpyDeviceToHost));
    if (sens64) HIPCHK(hipMemcpy(sens64, d64.p, d64.bytes, hipMemcpyDeviceToHost));
    return MLEM_OK;
}

// ---------------------------------------------------------------- recon
int mlem_recon_create(const mlem_matrix *mc, const int32_t *g, const float *sens, const float *f0,
                      const mlem_opts *opts, mlem_recon **out) {
    ARGCHK(out, "mlem_recon_create: null out");
    *out = nullptr;
    ARGCHK(mc && g && f0, "mlem_recon_create: null matrix, counts or image");
    mlem_matrix *m = const_cast<mlem_matrix *>(mc);  // only the partition cache changes
    mlem_opts o;
    if (opts)
        o = *opts;
    else
        mlem_opts_default(&o);
    ARGCHK(o.spmv <= 1, "mlem_recon_create: spmv must be 0 (merge-based) or 1 (CSR-vector)");
    ARGCHK(o.using_trans <= 1, "mlem_recon_create: using_trans must be 0 (transpose) or 1 (scatter)");
    ARGCHK(o.secsize_fp >= 1 && o.secsize_bp >= 1, "mlem_recon_create: section sizes must be >= 1");
    ARGCHK(o.fp_lanes == 0 || valid_lanes(o.fp_lanes), "mlem_recon_create: fp_lanes must be 0,4,8,16,32,64");
    ARGCHK(o.bp_lanes == 0 || valid_lanes(o.bp_lanes), "mlem_recon_create: bp_lanes must be 0,4,8,16,32,64");
    ARGCHK(o.bp_spmv <= 2, "mlem_recon_create: bp_spmv must be 0 (as spmv), 1 (merge) or 2 (CSR-vector)");
    if (o.using_trans == 0 && !m->has_T && !m->tiles_only)
        return set_err(MLEM_ERR_INVALID_ARGUMENT,
                       "mlem_recon_create: using_trans == 0 needs the transpose (upload it or build it)");
    if (m->tiles_only) {
        ARGCHK(o.spmv == 0 && o.using_trans == 0 && o.bp_spmv <= 1,
               "mlem_recon_create: a tiles-only matrix takes merge-based FP and BP on A^T");
        ARGCHK(m->tile_bc == 1, "mlem_recon_create: the matrix holds the tiles of batches of %u per pass", m->tile_bc);
        ARGCHK((o.tile_width == 0 || o.tile_width == m->tile_W) && (o.tile_rows == 0 || o.tile_rows == m->tile_R) &&
                   (o.values == 0 || o.values == m->tile_vt),
               "mlem_recon_create: the matrix holds tiles of width %u and %u rows with values %u", m->tile_W,
               m->tile_R, m->tile_vt);
        o.tile_width = m->tile_W;
        o.tile_rows = m->tile_R;
        o.values = m->tile_vt;
    }
    HIPCHK(hipSetDevice(m->device));
    auto r = std::make_unique<mlem_recon>();
    r->m = m;
    r->opts = o;
    r->opts.secsize_fp = clamp_secsize(o.secsize_fp);
    r->opts.secsize_bp = clamp_secsize(o.secsize_bp);
    r->lanes_fp = o.fp_lanes ? o.fp_lanes : auto_lanes(m->rows, m->nnz);
    r->lanes_bp = o.bp_lanes ? o.bp_lanes : auto_lanes(m->cols, m->nnz);
    r->bp_merge = o.bp_spmv == 0 ? o.spmv == 0 : o.bp_spmv == 1;
    HIPCHK(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
    struct StreamGuard {
        mlem_recon *r;
        bool armed = true;
        ~StreamGuard() {
            if (armed && r->stream) (void)hipStreamDestroy(r->stream);
        }
    } guard{r.get()};
    // f and ratio are the gathered vectors of FP / BP: padded with zeros to
    // whole segments of the blocked kernels.
    auto padded = [](size_t n) { return (n + kTileXPad - 1) / kTileXPad * kTileXPad + kTileXPad; };
    HIPCHK(r->g.alloc(sizeof(int32_t) * std::max<size_t>(1, m->rows)));
    HIPCHK(r->ratio.alloc(sizeof(float) * padded(m->rows)));
    HIPCHK(r->f.alloc(sizeof(float) * padded(m->cols)));
    HIPCHK(launch_fill(r->ratio.as<float>(), 0.0f, padded(m->rows), r->stream));
    HIPCHK(launch_fill(r->f.as<float>(), 0.0f, padded(m->cols), r->stream));
    HIPCHK(r->sens.alloc(sizeof(float) * m->cols));
    if (m->rows) HIPCHK(hipMemcpyAsync(r->g.p, g, sizeof(int32_t) * m->rows, hipMemcpyHostToDevice, r->stream));
    HIPCHK(hipMemcpyAsync(r->f.p, f0, sizeof(float) * m->cols, hipMemcpyHostToDevice, r->stream));
    if (sens) {
        HIPCHK(hipMemcpyAsync(r->sens.p, sens, sizeof(float) * m->cols, hipMemcpyHostToDevice, r->stream));
    } else {
        if (int st = sens_on_device(m, r->sens.as<float>(), nullptr, r->stream)) return st;
    }
    if (o.using_trans == 1) {
        HIPCHK(r->bp.alloc(sizeof(float) * m->cols));
        HIPCHK(launch_fill(r->bp.as<float>(), 0.0f, m->cols, r->stream));
    }
    ARGCHK(o.tile_width <= 1 || o.tile_width == 1024 || o.tile_width == 2048,
           "mlem_recon_create: tile_width must be 0 (auto), 1 (off), 1024 or 2048");
    ARGCHK(o.values <= VAL_UF16, "mlem_recon_create: values must be 0 (f32), 1 (bf16), 2 (fp16) or 3 (uf16)");
    ARGCHK(o.values == VAL_F32 || (o.spmv == 0 && o.using_trans == 0 && r->bp_merge),
           "mlem_recon_create: 16-bit values need merge-based FP and BP on A^T (the blocked kernels)");
    ARGCHK(o.tile_rows == 0 || valid_tile_rows(o.tile_rows),
           "mlem_recon_create: tile_rows must be 0 (auto), 512, 1024, 2048, 4096, 8192 or 16384");
    if (o.spmv == 0)
        if (int st = setup_merge_side(m, 0, o, r->opts.secsize_fp, &r->tfp, &r->wfp, &r->part_fp, &r->pfp,
                                      &r->scratch_fp))
            return st;
    if (o.using_trans == 0 && r->bp_merge)
        if (int st = setup_merge_side(m, 1, o, r->opts.secsize_bp, &r->tbp, &r->wbp, &r->part_bp, &r->pbp,
                                      &r->scratch_bp))
            return st;
    if (o.values != VAL_F32 && r->tbp)  // the quantized matrix's own sensitivity
        HIPCHK(hipMemcpyAsync(r->sens.p, r->tbp->sens.p, sizeof(float) * m->cols, hipMemcpyDeviceToDevice,
                              r->stream));
    HIPCHK(r->llpart.alloc(sizeof(double) * std::max<uint32_t>(1, loglik_blocks(m->rows))));
    HIPCHK(hipStreamSynchronize(r->stream));
    guard.armed = false;
    m->users.fetch_add(1);
    *out = r.release();
    return MLEM_OK;
}

int mlem_recon_iterate(mlem_recon *r, uint32_t iters) {
    ARGCHK(r, "mlem_recon_iterate: null handle");
    HIPCHK(hipSetDevice(r->m->device));
    if (r->opts.use_graph) {
        if (!r->graph_exec) {
            HIPCHK(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = enqueue_iteration(r, nullptr);
            hipGraph_t gr = nullptr;
            hipError_t e2 = hipStreamEndCapture(r->stream, &gr);
            if (e != hipSuccess) {
                if (gr) (void)hipGraphDestroy(gr);
                return hip_err(e, "enqueue_iteration (capture)", __FILE_NAME__, __LINE__);
            }
            HIPCHK(e2);
            r->graph = gr;
            HIPCHK(hipGraphInstantiate(&r->graph_exec, r->graph, nullptr, nullptr, 0));
        }
        for (uint32_t it = 0; it < iters; ++it) HIPCHK(hipGraphLaunch(r->graph_exec, r->stream));
        return MLEM_OK;
    }
    for (uint32_t it = 0; it < iters; ++it) {
        IterEvents *ev = nullptr;
        if (r->opts.profile) {
            if (r->events_used == r->events.size()) {
                IterEvents x;
                for (auto &e : x.e) HIPCHK(hipEventCreate(&e));
                r->events.push_back(x);
            }
            ev = &r->events[r->events_used++];
        }
        hipError_t e = enqueue_iteration(r, ev);
        if (e != hipSuccess) return hip_err(e, "enqueue_iteration", __FILE_NAME__, __LINE__);
    }
    return MLEM_OK;
}

int mlem_recon_sync(mlem_recon *r) {
    ARGCHK(r, "mlem_recon_sync: null handle");
    HIPCHK(hipSetDevice(r->m->device));
    HIPCHK(hipStreamSynchronize(r->stream));
    return collect_times(r);
}

int mlem_recon_get_image(mlem_recon *r, float *f_out) {
    ARGCHK(r && f_out, "mlem_recon_get_image: null argument");
    HIPCHK(hipSetDevice(r->m->device));
    HIPCHK(hipMemcpyAsync(f_out, r->f.p, sizeof(float) * r->m->cols, hipMemcpyDeviceToHost, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    return collect_times(r);
}

int mlem_recon_set_image(mlem_recon *r, const float *f) {
    ARGCHK(r && f, "mlem_recon_set_image: null argument");
    HIPCHK(hipSetDevice(r->m->device));
    HIPCHK(hipMemcpyAsync(r->f.p, f, sizeof(float) * r->m->cols, hipMemcpyHostToDevice, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    return MLEM_OK;
}

int mlem_recon_kernel_times(mlem_recon *r, double *fp_ms, double *bp_ms, uint32_t *iters_timed) {
    ARGCHK(r, "mlem_recon_kernel_times: null handle");
    if (int st = collect_times(r)) return st;
    if (fp_ms) *fp_ms = r->phase_ms[0] + r->phase_ms[1];
    if (bp_ms) *bp_ms = r->phase_ms[2] + r->phase_ms[3];
    if (iters_timed) *iters_timed = r->timed;
    return MLEM_OK;
}

int mlem_recon_phase_times(mlem_recon *r, double ms[4], uint32_t *iters_timed) {
    ARGCHK(r && ms, "mlem_recon_phase_times: null argument");
    if (int st = collect_times(r)) return st;
    for (int k = 0; k < 4; ++k) ms[k] = r->phase_ms[k];
    if (iters_timed) *iters_timed = r->timed;
    return MLEM_OK;
}

int mlem_recon_reset_times(mlem_recon *r) {
    ARGCHK(r, "mlem_recon_reset_times: null handle");
    if (int st = collect_times(r)) return st;
    for (double &v : r->phase_ms) v = 0;
    r->timed = 0;
    return MLEM_OK;
}

int mlem_recon_loglik(mlem_recon *r, double *ll) {
    ARGCHK(r && ll, "mlem_recon_loglik: null argument");
    HIPCHK(hipSetDevice(r->m->device));
    const mlem_matrix *m = r->m;
    const uint32_t nb = loglik_blocks(m->rows);
    if (m->tiles_only) {  // A f by the tiled forward projection
        DevBuf y;
        HIPCHK(y.alloc(sizeof(float) * std::max<size_t>(1, m->rows)));
        EpiArgs ea{};
        ea.out = y.as<float>();
        HIPCHK(launch_tiled_partials(r->tfp->plan, r->tfp->view, r->wfp->view, r->f.as<float>(), r->part_fp.as<float>(),
                                     1, 0, r->stream));
        HIPCHK(launch_tiled_reduce(r->tfp->plan, r->tfp->view, r->wfp->view, r->part_fp.as<float>(), 1, EPI_PLAIN, ea,
                                   r->stream));
        HIPCHK(launch_loglik_y(y.as<float>(), r->g.as<int32_t>(), m->rows, r->llpart.as<double>(), r->stream));
        HIPCHK(hipStreamSynchronize(r->stream));
    } else {
        HIPCHK(launch_loglik(m->rowptr.as<uint64_t>(), m->col.as<uint32_t>(), m->val.as<float>(), r->f.as<float>(),
                             r->g.as<int32_t>(), m->rows, r->llpart.as<double>(), r->stream));
    }
    std::vector<double> part(nb);
    if (nb) HIPCHK(hipMemcpyAsync(part.data(), r->llpart.p, sizeof(double) * nb, hipMemcpyDeviceToHost, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    double s = 0.0;
    for (double v : part) s += v;
    *ll = s;
    return MLEM_OK;
}

void *mlem_recon_stream(mlem_recon *r) { return r ? r->stream : nullptr; }

int mlem_recon_destroy(mlem_recon *r) {
    if (!r) return MLEM_OK;
    (void)hipSetDevice(r->m->device);
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    destroy_graph(r);
    for (auto &x : r->events)
        for (auto &e : x.e) (void)hipEventDestroy(e);
    if (r->stream) (void)hipStreamDestroy(r->stream);
    r->m->users.fetch_sub(1);
    delete r;
    return MLEM_OK;
}

int mlem_run(const mlem_matrix *m, const int32_t *g, const float *sens, const float *f0, float *f_out, uint32_t iters,
             const mlem_opts *opts) {
    ARGCHK(m && g && f0 && f_out, "mlem_run: null argument");
    mlem_recon *r = nullptr;
    if (int st = mlem_recon_create(m, g, sens, f0, opts, &r)) return st;
    int st = mlem_recon_iterate(r, iters);
    if (st == MLEM_OK) st = mlem_recon_get_image(r, f_out);
    if (st != MLEM_OK) {
        std::string msg = g_err;
        mlem_recon_destroy(r);
        g_err = msg;
        return st;
    }
    return mlem_recon_destroy(r);
}

}  // extern "C"
