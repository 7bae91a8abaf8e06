// Batches (many sinograms over one resident matrix, mlem_batch_*) and the
// tile-shape autotuner (mlem_autotune).
#include "abi_internal.hpp"

using namespace mlem;
using namespace mlem::abi;

namespace {

// One iteration of the passes marked in `passes` on the batch's stream.
int enqueue_batch_iteration(mlem_batch *b, const std::vector<uint8_t> &passes) {
    hipStream_t s = b->stream;
    const uint32_t BC = b->BC;
    for (uint32_t p = 0; p < b->passes; ++p) {
        if (!passes[p] || !b->pass[p]) continue;
        float *f = b->pass[p]->f.as<float>();
        float *ratio = b->pass[p]->ratio.as<float>();
        EpiArgs fa{};
        fa.out = ratio;
        fa.g = b->pass[p]->g.as<int32_t>();
        HIPCHK(launch_tiled_partials(b->tfp->plan, b->tfp->view, b->wfp->view, f, b->part_fp.as<float>(), BC, b->nt,
                                     s));
        HIPCHK(launch_tiled_reduce(b->tfp->plan, b->tfp->view, b->wfp->view, b->part_fp.as<float>(), BC, EPI_RATIO, fa,
                                   s));
        EpiArgs ba{};
        ba.norm = b->sens.as<float>();
        ba.f = f;
        ba.active = b->active.as<uint32_t>() + (size_t)p * BC;
        HIPCHK(launch_tiled_partials(b->tbp->plan, b->tbp->view, b->wbp->view, ratio, b->part_bp.as<float>(), BC,
                                     b->nt, s));
        HIPCHK(launch_tiled_reduce(b->tbp->plan, b->tbp->view, b->wbp->view, b->part_bp.as<float>(), BC, EPI_UPDATE,
                                   ba, s));
    }
    return MLEM_OK;
}

}  // namespace

namespace mlem {
namespace abi {

int batch_set_async(mlem_batch *b, uint32_t slot, const int32_t *g, const float *f0, uint64_t *allocated) {
    if (allocated) *allocated = 0;
    ARGCHK(b && g && f0, "mlem_batch_set: null argument");
    ARGCHK(slot < b->slots, "mlem_batch_set: slot %u >= %u", slot, b->slots);
    HIPCHK(hipSetDevice(b->m->device));
    const uint32_t p = slot / b->BC, q = slot % b->BC;
    const mlem_matrix *m = b->m;
    hipStream_t s = b->stream;
    if (!b->pass[p]) {  // first slot of this pass: its vectors, zero-filled (padding included)
        auto np = std::make_unique<BatchPass>();
        HIPCHK(np->g.alloc(sizeof(int32_t) * b->rows_x * b->BC));
        HIPCHK(np->ratio.alloc(sizeof(float) * b->rows_x * b->BC));
        HIPCHK(np->f.alloc(sizeof(float) * b->cols_x * b->BC));
        HIPCHK(launch_fill_u32(np->g.as<uint32_t>(), 0u, b->rows_x * b->BC, s));
        HIPCHK(launch_fill(np->ratio.as<float>(), 0.0f, b->rows_x * b->BC, s));
        HIPCHK(launch_fill(np->f.as<float>(), 0.0f, b->cols_x * b->BC, s));
        if (allocated) *allocated = np->g.bytes + np->ratio.bytes + np->f.bytes;
        b->pass[p] = std::move(np);
    }
    BatchPass &bp = *b->pass[p];
    // One staging vector serves both copies: the stream orders the second
    // copy after the first scatter.
    HIPCHK(hipMemcpyAsync(b->stage.p, g, sizeof(int32_t) * m->rows, hipMemcpyHostToDevice, s));
    if (b->row_order)  // the tiles hold A's rows in the matrix's row order: so do the counts
        HIPCHK(launch_ordered_put(bp.g.as<int32_t>() + q, b->BC, b->stage.p, b->row_order, m->rows, s));
    else
        HIPCHK(launch_strided_put(bp.g.as<int32_t>() + q, b->BC, b->stage.p, m->rows, s));
    HIPCHK(hipMemcpyAsync(b->stage.p, f0, sizeof(float) * m->cols, hipMemcpyHostToDevice, s));
    HIPCHK(launch_strided_put(bp.f.as<float>() + q, b->BC, b->stage.p, m->cols, s));
    HIPCHK(launch_fill_u32(b->active.as<uint32_t>() + slot, 1u, 1, s));
    b->active_host[slot] = 1;
    return MLEM_OK;
}

int batch_get_image_async(mlem_batch *b, uint32_t slot, float *f_out) {
    ARGCHK(b && f_out, "mlem_batch_get_image: null argument");
    ARGCHK(slot < b->slots, "mlem_batch_get_image: slot %u >= %u", slot, b->slots);
    const uint32_t p = slot / b->BC, q = slot % b->BC;
    ARGCHK(b->pass[p], "mlem_batch_get_image: slot %u was never set", slot);
    HIPCHK(hipSetDevice(b->m->device));
    HIPCHK(launch_strided_get(b->stage.p, b->pass[p]->f.as<float>() + q, b->BC, b->m->cols, b->stream));
    HIPCHK(hipMemcpyAsync(f_out, b->stage.p, sizeof(float) * b->m->cols, hipMemcpyDeviceToHost, b->stream));
    return MLEM_OK;
}

int batch_deactivate_async(mlem_batch *b, uint32_t slot) {
    ARGCHK(b, "mlem_batch_deactivate: null handle");
    ARGCHK(slot < b->slots, "mlem_batch_deactivate: slot %u >= %u", slot, b->slots);
    HIPCHK(hipSetDevice(b->m->device));
    b->active_host[slot] = 0;
    HIPCHK(launch_fill_u32(b->active.as<uint32_t>() + slot, 0u, 1, b->stream));
    return MLEM_OK;
}

int batch_move_slot_async(mlem_batch *b, uint32_t from, uint32_t to) {
    ARGCHK(b && from < b->slots && to < b->slots && from != to, "batch_move_slot: bad slots %u -> %u", from, to);
    const uint32_t BC = b->BC, pf = from / BC, qf = from % BC, pt = to / BC, qt = to % BC;
    ARGCHK(b->pass[pf] && b->pass[pt] && b->active_host[from] && !b->active_host[to],
           "batch_move_slot: slot %u is not active or slot %u is not free in an allocated pass", from, to);
    HIPCHK(hipSetDevice(b->m->device));
    hipStream_t s = b->stream;
    const uint32_t rows = b->m->rows, cols = b->m->cols;
    // The counts in the positions the tiles hold them (row order or not), and the image; the ratios are
    // recomputed from them by the next iteration.
    HIPCHK(launch_strided_get(b->stage.p, b->pass[pf]->g.as<int32_t>() + qf, BC, rows, s));
    HIPCHK(launch_strided_put(b->pass[pt]->g.as<int32_t>() + qt, BC, b->stage.p, rows, s));
    HIPCHK(launch_strided_get(b->stage.p, b->pass[pf]->f.as<float>() + qf, BC, cols, s));
    HIPCHK(launch_strided_put(b->pass[pt]->f.as<float>() + qt, BC, b->stage.p, cols, s));
    HIPCHK(launch_fill_u32(b->active.as<uint32_t>() + to, 1u, 1, s));
    HIPCHK(launch_fill_u32(b->active.as<uint32_t>() + from, 0u, 1, s));
    b->active_host[to] = 1;
    b->active_host[from] = 0;
    return MLEM_OK;
}

uint64_t batch_release_pass(mlem_batch *b, uint32_t p) {
    if (!b || p >= b->passes || !b->pass[p]) return 0;
    for (uint32_t q = 0; q < b->BC; ++q)
        if (b->active_host[(size_t)p * b->BC + q]) return 0;
    if (hipSetDevice(b->m->device) != hipSuccess || hipStreamSynchronize(b->stream) != hipSuccess) return 0;
    const uint64_t bytes = b->pass[p]->g.bytes + b->pass[p]->ratio.bytes + b->pass[p]->f.bytes;
    b->pass[p].reset();
    return bytes;
}

}  // namespace abi
}  // namespace mlem

extern "C" {

// -------------------------------------------------------------- autotuner
// Times the blocked kernels of both sides over a few tile shapes and
// workgroup counts and reports the fastest.  Candidate tiles that did not
// exist before and were not chosen are freed again.
int mlem_autotune(mlem_matrix *m, uint32_t per_pass, uint32_t values, mlem_tune_result *out) {
    ARGCHK(m && out, "mlem_autotune: null argument");
    std::memset(out, 0, sizeof(*out));
    const uint32_t BC = per_pass ? per_pass : 1u;
    ARGCHK(BC == 1 || BC == 4 || BC == 8, "mlem_autotune: per_pass must be 1, 4 or 8");
    ARGCHK(values <= VAL_UF16, "mlem_autotune: values must be 0 (f32), 1 (bf16), 2 (fp16) or 3 (uf16)");
    ARGCHK(m->has_T, "mlem_autotune: the transpose must be resident");
    HIPCHK(hipSetDevice(m->device));
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    std::vector<std::pair<uint32_t, uint32_t>> shapes;  // (W, R)
    std::vector<uint32_t> groups;                       // per CU
    std::vector<uint32_t> threads = {0};                // work-items per workgroup (several right-hand sides)
    if (BC > 1) threads = {256, 512};  // 128 was slower at every shape timed on the 3D scanner matrix
    if (BC == 1) {
        shapes = {{1024, 1024}, {1024, 2048}, {1024, 4096}, {2048, 1024}, {2048, 2048}};
        groups = {16, 24, 32};
    } else if (BC == 4) {
        shapes = {{512, 1024}, {512, 2048}, {1024, 1024}};
        groups = {8, 16};
    } else {
        shapes = {{512, 512}, {512, 1024}};
        groups = {8, 16};
    }
    std::set<mlem_matrix::TileKey> before;
    {
        std::lock_guard<std::mutex> lk(m->mu);
        for (const auto &t : m->tiles) before.insert(t.first);
    }
    const size_t rows_x = ((size_t)m->rows + kTileXPad - 1) / kTileXPad * kTileXPad + kTileXPad;
    const size_t cols_x = ((size_t)m->cols + kTileXPad - 1) / kTileXPad * kTileXPad + kTileXPad;
    DevBuf xf, xb, yf, yb;
    HIPCHK(xf.alloc(sizeof(float) * cols_x * BC));
    HIPCHK(xb.alloc(sizeof(float) * rows_x * BC));
    HIPCHK(yf.alloc(sizeof(float) * rows_x * BC));
    HIPCHK(yb.alloc(sizeof(float) * cols_x * BC));
    hipStream_t s = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    };
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        cleanup();
        return set_err(MLEM_ERR_HIP, "mlem_autotune: event creation failed");
    }
    if (launch_fill(xf.as<float>(), 1.0f, cols_x * BC, s) != hipSuccess ||
        launch_fill(xb.as<float>(), 1.0f, rows_x * BC, s) != hipSuccess) {
        cleanup();
        return set_err(MLEM_ERR_HIP, "mlem_autotune: fill failed");
    }
    double best = 1e300;
    uint32_t bestW = 0, bestR = 0, bestG = 0, bestT = 0, tried = 0;
    double best_fp = 0, best_bp = 0;
    // One side's time per launch (partials + reduce), mean of three after a warm-up.
    auto time_side = [&](const Tiled *t, const TileWorkDev *w, DevBuf &part, const float *x, float *y, uint32_t nt,
                         double *ms) {
        EpiArgs ea{};
        ea.out = y;
        for (int rep = 0; rep < 4; ++rep) {
            if (rep == 1 && hipEventRecord(e0, s) != hipSuccess) return hipErrorUnknown;
            hipError_t e = launch_tiled_partials(t->plan, t->view, w->view, x, part.as<float>(), BC, nt, s);
            if (e == hipSuccess) e = launch_tiled_reduce(t->plan, t->view, w->view, part.as<float>(), BC, EPI_PLAIN, ea, s);
            if (e != hipSuccess) return e;
        }
        hipError_t e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float f = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&f, e0, e1);
        *ms = f / 3.0;
        return e;
    };
    for (const auto &sh : shapes) {
        for (uint32_t g : groups) {
            const Tiled *tf = nullptr, *tb = nullptr;
            const TileWorkDev *wf = nullptr, *wb = nullptr;
            DevBuf pf, pb;
            int st = setup_tiled(m, 0, sh.first, sh.second, values, BC, g * (uint32_t)cus, &tf, &wf, &pf);
            if (st == MLEM_OK) st = setup_tiled(m, 1, sh.first, sh.second, values, BC, g * (uint32_t)cus, &tb, &wb, &pb);
            if (st == MLEM_ERR_UNSUPPORTED || st == MLEM_ERR_OUT_OF_MEMORY) {
                (void)hipGetLastError();
                continue;  // this shape does not suit or fit the matrix
            }
            if (st != MLEM_OK) {
                cleanup();
                return st;
            }
            for (uint32_t nt : threads) {
                double fp = 0, bp = 0;
                hipError_t e = time_side(tf, wf, pf, xf.as<float>(), yf.as<float>(), nt, &fp);
                if (e == hipSuccess) e = time_side(tb, wb, pb, xb.as<float>(), yb.as<float>(), nt, &bp);
                if (e != hipSuccess) {
                    cleanup();
                    return hip_err(e, "mlem_autotune: timing", __FILE_NAME__, __LINE__);
                }
                ++tried;
                if (fp + bp < best) {
                    best = fp + bp;
                    bestW = sh.first;
                    bestR = sh.second;
                    bestG = g;
                    bestT = nt;
                    best_fp = fp;
                    best_bp = bp;
                }
            }
        }
    }
    cleanup();
    {  // free the candidate tiles that were built here and not chosen
        std::lock_guard<std::mutex> lk(m->mu);
        for (auto it = m->tiles.begin(); it != m->tiles.end();) {
            const auto &k = it->first;
            const bool chosen = std::get<1>(k) == bestW && std::get<2>(k) == bestR && std::get<3>(k) == values &&
                                std::get<4>(k) == (BC > 1);
            it = (!before.count(k) && !chosen) ? m->tiles.erase(it) : std::next(it);
        }
    }
    if (!tried) return set_err(MLEM_ERR_UNSUPPORTED, "mlem_autotune: no tile shape suits or fits this matrix");
    out->tile_width = bestW;
    out->tile_rows = bestR;
    out->tile_groups = bestG * (uint32_t)cus;
    out->tile_threads = bestT;
    out->candidates = tried;
    out->fp_ms = best_fp;
    out->bp_ms = best_bp;
    return MLEM_OK;
}

// ---------------------------------------------------------------- batches
void mlem_batch_opts_default(mlem_batch_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->per_pass = 4;
}

int mlem_batch_create(mlem_matrix *m, uint32_t slots, const mlem_batch_opts *opts, mlem_batch **out) {
    ARGCHK(out, "mlem_batch_create: null out");
    *out = nullptr;
    ARGCHK(m && slots > 0, "mlem_batch_create: null matrix or no slots");
    mlem_batch_opts o;
    if (opts)
        o = *opts;
    else
        mlem_batch_opts_default(&o);
    const uint32_t BC = o.per_pass ? o.per_pass : 4u;
    ARGCHK(BC == 1 || BC == 4 || BC == 8, "mlem_batch_create: per_pass must be 1, 4 or 8");
    ARGCHK(o.values <= VAL_UF16, "mlem_batch_create: values must be 0 (f32), 1 (bf16), 2 (fp16) or 3 (uf16)");
    ARGCHK(m->has_T || m->tiles_only, "mlem_batch_create: the transpose must be resident (upload it or build it)");
    if (m->tiles_only) {
        ARGCHK(BC == m->tile_bc && (o.tile_width == 0 || o.tile_width == m->tile_W) &&
                   (o.tile_rows == 0 || o.tile_rows == m->tile_R) && (o.values == 0 || o.values == m->tile_vt),
               "mlem_batch_create: the matrix holds tiles of width %u and %u rows with values %u for %u per pass",
               m->tile_W, m->tile_R, m->tile_vt, m->tile_bc);
        o.tile_width = m->tile_W;
        o.tile_rows = m->tile_R;
        o.values = m->tile_vt;
    }
    // Defaults: 1024 x 2048 for one per pass, 1024 x 1024 for 4 (3.28x the
    // single FP32 path at C2 against 3.17x for 512 x 2048), 512 x 1024 for 8;
    // 512 rows for small matrices (default_tile_rows).
    const uint32_t W = o.tile_width ? o.tile_width : (BC == 8 ? 512u : 1024u);
    ARGCHK(valid_tile_config(W, BC), "mlem_batch_create: no blocked kernel for width %u with %u per pass", W, BC);
    ARGCHK(o.tile_rows == 0 || valid_tile_rows(o.tile_rows),
           "mlem_batch_create: tile_rows must be 512 ... 16384 (a power of two)");
    ARGCHK(valid_stream_threads(o.tile_threads), "mlem_batch_create: tile_threads must be 0, 128, 256 or 512");
    HIPCHK(hipSetDevice(m->device));
    const uint32_t R = o.tile_rows ? o.tile_rows : default_tile_rows(m->nnz, o.tile_groups, BC);
    auto b = std::make_unique<mlem_batch>();
    b->m = m;
    b->slots = slots;
    b->BC = BC;
    b->nt = o.tile_threads;
    b->use_graph = o.use_graph != 0;
    b->passes = (slots + BC - 1) / BC;
    b->rows_x = (m->rows + kTileXPad - 1) / kTileXPad * kTileXPad + kTileXPad;
    b->cols_x = (m->cols + kTileXPad - 1) / kTileXPad * kTileXPad + kTileXPad;
    HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    int cus = 256, dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const uint32_t groups = o.tile_groups ? o.tile_groups : (BC == 1 ? 24u : 8u) * (uint32_t)std::max(1, cus);
    if (int st = setup_tiled(m, 0, W, R, o.values, BC, groups, &b->tfp, &b->wfp, &b->part_fp)) return st;
    if (int st = setup_tiled(m, 1, W, R, o.values, BC, groups, &b->tbp, &b->wbp, &b->part_bp)) return st;
    if (b->tfp->ordered != b->tbp->ordered)  // a row order set from another thread between the two builds
        return set_err(MLEM_ERR_INVALID_ARGUMENT, "mlem_batch_create: the row order changed while the tiles were built");
    if (b->tfp->ordered) b->row_order = m->row_order.as<uint32_t>();
    const size_t per = (size_t)BC * b->passes;
    b->pass.resize(b->passes);  // passes are allocated when first used (batch_set_async)
    HIPCHK(b->sens.alloc(sizeof(float) * m->cols));
    HIPCHK(b->active.alloc(sizeof(uint32_t) * per));
    HIPCHK(b->stage.alloc(sizeof(float) * std::max<size_t>(1, std::max(m->rows, m->cols))));
    HIPCHK(launch_fill_u32(b->active.as<uint32_t>(), 0u, per, b->stream));
    b->active_host.assign(per, 0);
    if (o.values != VAL_F32)  // the quantized matrix's own sensitivity
        HIPCHK(hipMemcpyAsync(b->sens.p, b->tbp->sens.p, sizeof(float) * m->cols, hipMemcpyDeviceToDevice,
                              b->stream));
    else if (int st = sens_on_device(m, b->sens.as<float>(), nullptr, b->stream))
        return st;
    HIPCHK(hipStreamSynchronize(b->stream));
    m->users.fetch_add(1);
    *out = b.release();
    return MLEM_OK;
}

int mlem_batch_set(mlem_batch *b, uint32_t slot, const int32_t *g, const float *f0) {
    if (int st = batch_set_async(b, slot, g, f0, nullptr)) return st;
    HIPCHK(hipStreamSynchronize(b->stream));  // g and f0 may be freed after the return
    return MLEM_OK;
}

int mlem_batch_deactivate(mlem_batch *b, uint32_t slot) { return batch_deactivate_async(b, slot); }

int mlem_batch_iterate(mlem_batch *b, uint32_t iters) {
    ARGCHK(b, "mlem_batch_iterate: null handle");
    HIPCHK(hipSetDevice(b->m->device));
    const uint32_t BC = b->BC;
    std::vector<uint8_t> passes(b->passes, 0);
    bool any = false;
    for (uint32_t p = 0; p < b->passes; ++p) {
        for (uint32_t q = 0; q < BC; ++q) passes[p] = passes[p] || b->active_host[(size_t)p * BC + q];
        any = any || passes[p];
    }
    if (!any || iters == 0) return MLEM_OK;
    if (!b->use_graph) {
        for (uint32_t it = 0; it < iters; ++it)
            if (int st = enqueue_batch_iteration(b, passes)) return st;
        return MLEM_OK;
    }
    if (!b->gexec || passes != b->graph_passes) {  // (re)capture for this set of passes
        if (b->gexec) {
            HIPCHK(hipStreamSynchronize(b->stream));
            HIPCHK(hipGraphExecDestroy(b->gexec));
            b->gexec = nullptr;
        }
        HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const int st = enqueue_batch_iteration(b, passes);
        hipGraph_t gr = nullptr;
        const hipError_t e2 = hipStreamEndCapture(b->stream, &gr);
        if (st != MLEM_OK) {
            if (gr) (void)hipGraphDestroy(gr);
            return st;
        }
        HIPCHK(e2);
        const hipError_t e3 = hipGraphInstantiate(&b->gexec, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        HIPCHK(e3);
        b->graph_passes = passes;
    }
    for (uint32_t it = 0; it < iters; ++it) HIPCHK(hipGraphLaunch(b->gexec, b->stream));
    return MLEM_OK;
}

int mlem_batch_sync(mlem_batch *b) {
    ARGCHK(b, "mlem_batch_sync: null handle");
    HIPCHK(hipSetDevice(b->m->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    return MLEM_OK;
}

int mlem_batch_get_image(mlem_batch *b, uint32_t slot, float *f_out) {
    if (int st = batch_get_image_async(b, slot, f_out)) return st;
    HIPCHK(hipStreamSynchronize(b->stream));
    return MLEM_OK;
}

int mlem_batch_info(const mlem_batch *b, uint32_t *slots, uint32_t *per_pass, uint32_t *passes,
                    uint64_t *device_bytes) {
    ARGCHK(b, "mlem_batch_info: null handle");
    if (slots) *slots = b->slots;
    if (per_pass) *per_pass = b->BC;
    if (passes) *passes = b->passes;
    if (device_bytes) *device_bytes = b->device_bytes();
    return MLEM_OK;
}

void *mlem_batch_stream(mlem_batch *b) { return b ? b->stream : nullptr; }

int mlem_batch_destroy(mlem_batch *b) {
    if (!b) return MLEM_OK;
    (void)hipSetDevice(b->m->device);
    b->m->users.fetch_sub(1);
    delete b;
    return MLEM_OK;
}

}  // extern "C"
