// Matrices resident as tiles only (mlem_matrix_create_tiled): the tiles of
// A and A^T built in chunks of whole row blocks, so that neither CSR copy has
// to stay on the device beside them.
#include "abi_internal.hpp"

using namespace mlem;
using namespace mlem::abi;

namespace {

// Tiles (one right-hand side; f32 or 16-bit values) of a matrix that is not
// resident as one CSR: given as chunks of whole row blocks, each a CSR of its own
// (rows r0 .. r0+rows_c, row pointers from 0) on the device.  Two passes
// over the same row blocks (the chunks may differ between them): the counts
// give every tile's size and so the offsets of the whole layout; then, the
// tiles allocated, each chunk fills its part.  The bytes are those that one
// build over the whole CSR writes (every tile is laid out on its own).
struct ChunkedTiles {
    TilePlan P;
    std::vector<uint64_t> size;  // padded entries per tile, then (allocate) offsets
    std::vector<uint32_t> nz;    // entries per tile
    Tiled *t = nullptr;

    ChunkedTiles(const TilePlan &p, Tiled *tiled) : P(p), size(p.ntiles + 1, 0), nz(p.ntiles, 0), t(tiled) {}

    // The chunk's own plan, its count into `scratch` (allocated here).
    int count_chunk(uint32_t r0, uint32_t rows_c, uint64_t nnz_c, const uint64_t *rp, const uint32_t *ci,
                    TilePlan *pc, DevBuf *scratch) {
        *pc = tile_plan(rows_c, P.cols, nnz_c, P.W, P.R, P.vt, P.quad);
        pc->unit_scale = P.unit_scale;
        HIPCHK(scratch->alloc(tile_count_scratch_bytes(*pc)));
        uint32_t fl = 0;
        HIPCHK(launch_tile_count(*pc, rp, ci, scratch->p, &fl, nullptr));
        if (fl) return set_err(MLEM_ERR_UNSUPPORTED, "tiling refused: a row is not sorted by column (rows from %u)", r0);
        return MLEM_OK;
    }

    int count(uint32_t r0, uint32_t rows_c, uint64_t nnz_c, const uint64_t *rp, const uint32_t *ci) {
        TilePlan pc;
        DevBuf scratch;
        if (int st = count_chunk(r0, rows_c, nnz_c, rp, ci, &pc, &scratch)) return st;
        const uint64_t t0 = (uint64_t)(r0 / P.R) * P.nseg;
        std::vector<uint64_t> off(pc.ntiles + 1);
        HIPCHK(hipMemcpy(off.data(), tile_count_offsets(pc, scratch.p), 8 * (pc.ntiles + 1), hipMemcpyDeviceToHost));
        if (pc.ntiles)
            HIPCHK(hipMemcpy(nz.data() + t0, tile_count_sizes(pc, scratch.p), 4 * pc.ntiles, hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i < pc.ntiles; ++i) size[t0 + i] = off[i + 1] - off[i];
        return MLEM_OK;
    }

    int allocate() {
        uint64_t run = 0;
        for (uint64_t i = 0; i <= P.ntiles; ++i) {
            const uint64_t n = i < P.ntiles ? size[i] : 0;
            size[i] = run;
            run += n;
        }
        P.cap = run + 8;  // as one build: a multiple of 8, at least one group
        t->plan = P;
        HIPCHK(t->mem.alloc(tiled_matrix_bytes(P)));
        t->view = tiled_matrix_view(P, t->mem.p);
        HIPCHK(hipMemcpy(t->view.tile_off, size.data(), 8 * (P.ntiles + 1), hipMemcpyHostToDevice));
        if (P.ntiles) HIPCHK(hipMemcpy(t->view.tile_nz, nz.data(), 4 * P.ntiles, hipMemcpyHostToDevice));
        return MLEM_OK;
    }

    int fill(uint32_t r0, uint32_t rows_c, uint64_t nnz_c, const uint64_t *rp, const uint32_t *ci, const float *va) {
        TilePlan pc;
        DevBuf scratch, local_off, fill_scratch;
        if (int st = count_chunk(r0, rows_c, nnz_c, rp, ci, &pc, &scratch)) return st;
        const uint64_t t0 = (uint64_t)(r0 / P.R) * P.nseg;
        HIPCHK(local_off.alloc(8 * (pc.ntiles + 1)));
        HIPCHK(fill_scratch.alloc(tile_fill_scratch_bytes(pc)));
        TiledMatrix v;
        v.tile_off = local_off.as<uint64_t>();
        v.tile_nz = t->view.tile_nz + t0;
        v.tdata = t->view.tdata + (uint64_t)(P.vb() + 2) * size[t0];
        v.scale = t->view.scale ? t->view.scale + r0 : nullptr;
        uint32_t fl = 0;
        HIPCHK(launch_tile_fill_into(pc, rp, ci, va, v, scratch.p, fill_scratch.p, &fl, nullptr));
        if (fl & 8u) return set_err(MLEM_ERR_INVALID_ARGUMENT, "uf16 values need a matrix without negative values");
        if (fl) return set_err(MLEM_ERR_FORMAT, "tiling: layout check failed (flags %u, rows from %u)", fl, r0);
        return MLEM_OK;
    }
};

}  // namespace

extern "C" {

int mlem_matrix_create_tiled(int device, uint32_t rows, uint32_t cols, uint64_t nnz, const uint64_t *rowptr,
                             const uint32_t *col, const float *val, uint32_t tile_width, uint32_t tile_rows,
                             uint32_t values, uint32_t per_pass, uint64_t chunk_bytes, mlem_matrix **out) {
    ARGCHK(out, "mlem_matrix_create_tiled: null out");
    *out = nullptr;
    ARGCHK(rowptr && (nnz == 0 || (col && val)), "mlem_matrix_create_tiled: null matrix array");
    ARGCHK(rows > 0 && cols > 0, "mlem_matrix_create_tiled: rows and cols must be > 0");
    if (int st = check_csr_host("mlem_matrix_create_tiled", rows, cols, nnz, rowptr, col)) return st;
    const uint32_t BC = per_pass ? per_pass : 1u;
    ARGCHK(BC == 1 || BC == 4 || BC == 8, "mlem_matrix_create_tiled: per_pass must be 0, 1, 4 or 8");
    const bool quad = BC > 1;  // the batches' layout
    const uint32_t W = tile_width ? tile_width : (BC == 8 ? 512u : 1024u);  // as mlem_batch_create
    ARGCHK(valid_tile_config(W, BC), "mlem_matrix_create_tiled: no blocked kernel for width %u with %u per pass", W, BC);
    ARGCHK(tile_rows == 0 || valid_tile_rows(tile_rows),
           "mlem_matrix_create_tiled: tile_rows must be 512 ... 16384 (a power of two)");
    ARGCHK(values <= VAL_UF16, "mlem_matrix_create_tiled: values must be 0 (f32), 1 (bf16), 2 (fp16) or 3 (uf16)");
    HIPCHK(hipSetDevice(device));
    // as a reconstruction over the resident CSR would choose
    const uint32_t R = tile_rows ? tile_rows : default_tile_rows(nnz, 0, BC);
    auto m = std::make_unique<mlem_matrix>();
    m->device = device;
    m->rows = rows;
    m->cols = cols;
    m->nnz = nnz;
    m->tiles_only = true;
    m->tile_W = W;
    m->tile_R = R;
    m->tile_vt = values;
    m->tile_bc = BC;
    const size_t margin = (size_t)1 << 30;
    auto avail = [&](size_t *b) -> int {
        size_t free_b = 0, total_b = 0;
        HIPCHK(hipMemGetInfo(&free_b, &total_b));
        *b = free_b > margin ? (free_b - margin) / 20 * 17 : 0;  // 85%: the free memory is not one piece
        if (chunk_bytes) *b = std::min<size_t>(*b, chunk_bytes);
        return MLEM_OK;
    };
    HIPCHK(m->sens32.alloc(sizeof(float) * cols));
    HIPCHK(m->sens64.alloc(sizeof(double) * cols));
    auto tT = std::make_unique<Tiled>();
    auto tA = std::make_unique<Tiled>();
    TilePlan pT = tile_plan(cols, rows, nnz, W, R, values, quad);
    pT.unit_scale = values != VAL_F32;  // A^T's tiles hold A's codes (decoded below, in place)
    ChunkedTiles bT(pT, tT.get());
    ChunkedTiles bA(tile_plan(rows, cols, nnz, W, R, values, quad), tA.get());
    DevBuf scaleA;  // 16-bit values: A's row scales (for the sensitivity)
    {
        // A on the device; A^T's tiles from slabs of A^T's rows (column
        // ranges of A, whole row blocks of A^T), the sensitivity from the
        // slabs' row sums (FP64, as for a resident transpose).
        DevBuf rp, ci, va, colptrT, srp;
        HIPCHK(rp.alloc(sizeof(uint64_t) * ((size_t)rows + 1)));
        HIPCHK(ci.alloc(sizeof(uint32_t) * std::max<uint64_t>(1, nnz)));
        HIPCHK(va.alloc(sizeof(float) * std::max<uint64_t>(1, nnz)));
        if (int st = upload_h2d(rp.p, rowptr, rp.bytes)) return st;
        if (nnz) {
            if (int st = upload_h2d(ci.p, col, sizeof(uint32_t) * nnz)) return st;
            if (int st = upload_h2d(va.p, val, sizeof(float) * nnz)) return st;
        }
        if (int st = check_columns_device("mlem_matrix_create_tiled", ci.as<uint32_t>(), nnz, cols)) return st;
        HIPCHK(colptrT.alloc(sizeof(uint64_t) * ((size_t)cols + 1)));
        HIPCHK(srp.alloc(sizeof(uint64_t) * ((size_t)cols + 1)));
        if (values != VAL_F32) {  // A's values become their decoded codes: what A^T's tiles store
            DevBuf bad;
            HIPCHK(scaleA.alloc(sizeof(float) * rows));
            HIPCHK(bad.alloc(sizeof(uint32_t)));
            HIPCHK(launch_fill(bad.as<float>(), 0.0f, 1, nullptr));
            HIPCHK(launch_quantize_in_place(rp.as<uint64_t>(), va.as<float>(), rows, values, scaleA.as<float>(),
                                            bad.as<uint32_t>(), nullptr));
            uint32_t fl = 0;
            HIPCHK(hipMemcpy(&fl, bad.p, sizeof(uint32_t), hipMemcpyDeviceToHost));
            if (fl & 8u)
                return set_err(MLEM_ERR_INVALID_ARGUMENT, "uf16 values need a matrix without negative values");
        }
        for (int pass = 0; pass < 2; ++pass) {
            if (pass == 1)
                if (int st = bT.allocate()) return st;
            int st_sink = MLEM_OK;
            RangeSink sink = [&](uint32_t c0, uint32_t c1, uint64_t off0, uint64_t n, const uint32_t *rowT,
                                 const float *valT) -> hipError_t {
                hipError_t e = launch_rebase(colptrT.as<uint64_t>() + c0, off0, c1 - c0, srp.as<uint64_t>(), nullptr);
                if (e != hipSuccess) return e;
                if (pass == 0) {
                    // FP64 column sums of A (of the quantized matrix: codes times row scales)
                    e = values == VAL_F32
                            ? launch_colsum_rows_f64(srp.as<uint64_t>(), valT, c1 - c0, m->sens64.as<double>() + c0,
                                                     m->sens32.as<float>() + c0, nullptr)
                            : launch_colsum_weighted_f64(srp.as<uint64_t>(), rowT, valT, scaleA.as<float>(), c1 - c0,
                                                         m->sens64.as<double>() + c0, m->sens32.as<float>() + c0,
                                                         nullptr);
                    if (e != hipSuccess) return e;
                    st_sink = bT.count(c0, c1 - c0, n, srp.as<uint64_t>(), rowT);
                } else {
                    st_sink = bT.fill(c0, c1 - c0, n, srp.as<uint64_t>(), rowT, valT);
                }
                return st_sink == MLEM_OK ? hipSuccess : hipErrorUnknown;
            };
            // Half of what is free for the ranges' sorts, the rest for the
            // slab's tiling scratch.
            size_t budget = 0;
            if (int st = avail(&budget)) return st;
            uint32_t nr = 0;
            const hipError_t e = transpose_ranges_to(rows, cols, nnz, rp.as<uint64_t>(), ci.as<uint32_t>(),
                                                     va.as<float>(), colptrT.as<uint64_t>(), nullptr, nullptr,
                                                     budget / 2, R, sink, &nr, nullptr);
            if (st_sink != MLEM_OK) return st_sink;
            if (e == hipErrorOutOfMemory)
                return set_err(MLEM_ERR_OUT_OF_MEMORY, "mlem_matrix_create_tiled: %u rows of A^T do not fit", R);
            HIPCHK(e);
            progress("A^T tiles %s over %u column ranges (%.1f GB of range scratch)", pass ? "filled" : "counted", nr,
                     budget / 2 / 1e9);
        }
        HIPCHK(hipDeviceSynchronize());
    }
    // A's tiles from row chunks of the host CSR (A's device copy is gone),
    // whole row blocks whose CSR and tiling scratch fit.
    const uint64_t nseg = bA.P.nseg;
    progress("A^T tiles %.1f GB; A's device copy freed", tT->mem.bytes / 1e9);
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1)
            if (int st = bA.allocate()) return st;
        size_t budget = 0;
        if (int st = avail(&budget)) return st;
        uint32_t chunks = 0, retries = 0;
        // Per nonzero: the chunk's CSR (8 bytes), the fill's column and row
        // scratch (4, with padding), 16-bit codes (2); per row: the
        // per-segment counts and positions.
        auto need = [&](uint32_t a, uint32_t b) {
            const uint64_t n = rowptr[b] - rowptr[a];
            return n * (values == VAL_F32 ? 16 : 18) + (uint64_t)(b - a) * (nseg * 8 + 8) + ((size_t)256 << 20);
        };
        // One chunk of whole row blocks [r0, r1): uploaded, counted or filled.
        auto chunk = [&](uint32_t r0, uint32_t r1) -> int {
            const uint32_t rc = r1 - r0;
            const uint64_t n = rowptr[r1] - rowptr[r0];
            std::vector<uint64_t> hrp(rc + 1);
            for (uint32_t i = 0; i <= rc; ++i) hrp[i] = rowptr[r0 + i] - rowptr[r0];
            DevBuf rp, ci, va;
            HIPCHK(rp.alloc(sizeof(uint64_t) * (rc + 1)));
            HIPCHK(ci.alloc(sizeof(uint32_t) * std::max<uint64_t>(1, n)));
            HIPCHK(hipMemcpy(rp.p, hrp.data(), rp.bytes, hipMemcpyHostToDevice));
            if (n)
                if (int st = upload_h2d(ci.p, col + rowptr[r0], sizeof(uint32_t) * n)) return st;
            if (pass == 0) return bA.count(r0, rc, n, rp.as<uint64_t>(), ci.as<uint32_t>());
            HIPCHK(va.alloc(sizeof(float) * std::max<uint64_t>(1, n)));
            if (n)
                if (int st = upload_h2d(va.p, val + rowptr[r0], sizeof(float) * n)) return st;
            return bA.fill(r0, rc, n, rp.as<uint64_t>(), ci.as<uint32_t>(), va.as<float>());
        };
        for (uint32_t r0 = 0; r0 < rows; ++chunks) {
            uint32_t r1 = std::min<uint64_t>((uint64_t)r0 + R, rows);
            while (r1 < rows) {
                const uint32_t nx = (uint32_t)std::min<uint64_t>((uint64_t)r1 + R, rows);
                if (need(r0, nx) > budget) break;
                r1 = nx;
            }
            // An allocation that fails although the estimate fit (the free
            // memory is not one piece): the same start with half the blocks.
            for (;;) {
                const int st = chunk(r0, r1);
                if (st == MLEM_OK) break;
                const uint32_t blocks = (r1 - r0 + R - 1) / R;
                if (st != MLEM_ERR_OUT_OF_MEMORY || blocks < 2) return st;
                (void)hipGetLastError();
                r1 = r0 + (blocks / 2) * R;
                ++retries;
            }
            r0 = r1;
        }
        progress("A tiles %s over %u row chunks (%u smaller after a failed allocation; %.1f GB of chunk scratch)",
                 pass ? "filled" : "counted", chunks, retries, budget / 1e9);
    }
    HIPCHK(hipDeviceSynchronize());
    progress("A tiles %.1f GB", tA->mem.bytes / 1e9);
    if (values != VAL_F32) {  // the quantized matrix's own sensitivity, as the resident path keeps it
        HIPCHK(tT->sens.alloc(sizeof(float) * cols));
        HIPCHK(hipMemcpy(tT->sens.p, m->sens32.p, sizeof(float) * cols, hipMemcpyDeviceToDevice));
    }
    m->tiles.emplace(std::make_tuple(0, W, R, values, quad, false), std::move(tA));
    m->tiles.emplace(std::make_tuple(1, W, R, values, quad, false), std::move(tT));
    *out = m.release();
    return MLEM_OK;
}

}  // extern "C"
