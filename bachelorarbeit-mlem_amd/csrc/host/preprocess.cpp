// Host preprocessing that feeds the MLEM loop: f32 column sums, the initial
// image, the stable counting-sort transpose and the nnz-balanced partition.
#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "host_common.hpp"

using mlem_host::fail;

extern "C" {

// A parallel memcpy on the library's OpenMP threads (which follow
// OMP_PROC_BIND / OMP_PLACES, unlike threads started elsewhere in a process
// whose initial thread OpenMP has bound to one core): page-aligned slices.
int mlem_host_copy(void *dst, const void *src, uint64_t bytes, int threads) {
    if (bytes && (!dst || !src)) return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_host_copy: null argument");
    if (bytes == 0) return MLEM_OK;
    const int nt = threads > 0 ? threads : omp_get_max_threads();
    const uint64_t per = ((bytes + (uint64_t)nt - 1) / (uint64_t)nt + 4095) & ~uint64_t(4095);  // >= 4096
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int t = 0; t < nt; ++t) {
        const uint64_t b = std::min<uint64_t>(bytes, per * (uint64_t)t), e = std::min<uint64_t>(bytes, per * (uint64_t)(t + 1));
        if (b < e) std::memcpy(static_cast<uint8_t *>(dst) + b, static_cast<const uint8_t *>(src) + b, e - b);
    }
    return MLEM_OK;
}

// calcColumnSums (main.cu:558-580).  The reference reduces thread-private
// float vectors with an OpenMP user-defined reduction, so its rounding
// depends on the thread count; this is the one-thread order: each column
// accumulates its entries in increasing row order.
int mlem_colsum_f32(const uint64_t *rowptr, const uint32_t *col, const float *val, uint32_t rows,
                    uint32_t cols, float *norm) {
    if (!rowptr || !norm || (rowptr[rows] && (!col || !val)))
        return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_colsum_f32: null argument");
    std::fill(norm, norm + cols, 0.0f);
    for (uint32_t r = 0; r < rows; ++r)
        for (uint64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            if (col[k] >= cols) return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_colsum_f32: column %u out of range", col[k]);
            norm[col[k]] += val[k];
        }
    return MLEM_OK;
}

// main.cu:83-97: serial float sum of the norm, int sum of the counts, then
// init = sum_g / sum_norm (int converted to float, float division).
int mlem_initial_image(const int32_t *g, uint32_t rows, const float *norm, uint32_t cols, float *f0,
                       float *init_out) {
    if ((rows && !g) || (cols && (!norm || !f0)))
        return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_initial_image: null argument");
    float sum_norm = 0.0f;
    for (uint32_t i = 0; i < cols; ++i) sum_norm += norm[i];
    // The reference sums into an int; wrap-around is made explicit with u32.
    uint32_t sum_g_bits = 0;
    for (uint32_t i = 0; i < rows; ++i) sum_g_bits += (uint32_t)g[i];
    int32_t sum_g = (int32_t)sum_g_bits;
    float init = (float)sum_g / sum_norm;
    for (uint32_t i = 0; i < cols; ++i) f0[i] = init;
    if (init_out) *init_out = init;
    return MLEM_OK;
}

// CSR(A) -> CSR(A^T) as a stable counting sort.
//
// sptrans_scanTrans_specialized (sptrans.h:397-545) splits the nnz into
// `procs` contiguous chunks (the first nnz % procs chunks one longer), counts
// per (chunk, column), prefix-sums the counts over chunks and over columns,
// and writes every entry to colptrT[c] + (entries of column c in earlier
// chunks) + (entries of column c earlier in its own chunk).  That places
// each column's entries in increasing nnz position, i.e. increasing row: the
// output is the stable counting sort of the entries by column and does not
// depend on procs.  This version keeps the same chunking but turns the
// per-chunk counts into running write cursors in place, so it needs only
// (chunks+1) x cols counters instead of sptrans' three nnz-long u64 arrays
// (csrRowUnroll, intra, wb_index).
//
// Difference kept on purpose: sptrans' block sums in scan_omp_specialzed use
// an `int` accumulator (std::accumulate(..., 0), sptrans.h:232), which wraps
// once 2048 consecutive columns hold >= 2^31 entries; the counts here are
// exact 64-bit values, so both agree whenever the reference is well defined.
int mlem_transpose(uint32_t rows, uint32_t cols, uint64_t nnz, const uint64_t *rowptr, const uint32_t *col,
                   const float *val, uint64_t *colptrT, uint32_t *rowT, float *valT, int threads) {
    if (!rowptr || !colptrT || (nnz && (!col || !rowT)))
        return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_transpose: null argument");
    if (rowptr[0] != 0 || rowptr[rows] != nnz)
        return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_transpose: rowptr must run from 0 to nnz");
    int procs = threads > 0 ? threads : omp_get_max_threads();
    if ((uint64_t)procs > nnz) procs = nnz > 0 ? (int)nnz : 1;
    const uint64_t size = nnz / (uint64_t)procs, left = nnz % (uint64_t)procs;
    auto chunk_begin = [&](uint64_t t) { return t < left ? t * size + t : t * size + left; };
    auto chunk_len = [&](uint64_t t) { return size + (t < left ? 1 : 0); };

    std::vector<uint64_t> cnt;
    try {
        cnt.assign((size_t)procs * cols, 0);
    } catch (...) {
        return fail(MLEM_ERR_OUT_OF_MEMORY, "mlem_transpose: cannot allocate %d x %u counters", procs, cols);
    }
    int bad = 0;
    // Pass 1 (sptrans.h:420-442): per-chunk column histogram.
#pragma omp parallel num_threads(procs) reduction(| : bad)
    {
        uint64_t t = (uint64_t)omp_get_thread_num();
        uint64_t *c = cnt.data() + t * cols;
        uint64_t b = chunk_begin(t), e = b + chunk_len(t);
        for (uint64_t k = b; k < e; ++k) {
            uint32_t cc = col[k];
            if (cc >= cols) { bad = 1; break; }
            c[cc]++;
        }
    }
    if (bad) return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_transpose: column index >= cols");
    // Column totals and the inclusive scan into colptrT (sptrans.h:444-457).
    colptrT[0] = 0;
    {
        std::vector<uint64_t> tot(cols, 0);
#pragma omp parallel for num_threads(procs) schedule(static)
        for (int64_t c = 0; c < (int64_t)cols; ++c) {
            uint64_t s = 0;
            for (int t = 0; t < procs; ++t) s += cnt[(size_t)t * cols + c];
            tot[c] = s;
        }
        uint64_t run = 0;
        for (uint32_t c = 0; c < cols; ++c) {
            run += tot[c];
            colptrT[c + 1] = run;
        }
    }
    // Counts -> write cursors: cursor[t][c] = colptrT[c] + sum_{t'<t} cnt[t'][c].
#pragma omp parallel for num_threads(procs) schedule(static)
    for (int64_t c = 0; c < (int64_t)cols; ++c) {
        uint64_t run = colptrT[c];
        for (int t = 0; t < procs; ++t) {
            uint64_t n = cnt[(size_t)t * cols + c];
            cnt[(size_t)t * cols + c] = run;
            run += n;
        }
    }
    // Scatter in nnz order within each chunk (sptrans.h:459-515); the row of
    // entry k is tracked by walking rowptr from the chunk's first row.
#pragma omp parallel num_threads(procs)
    {
        uint64_t t = (uint64_t)omp_get_thread_num();
        uint64_t *cur = cnt.data() + t * cols;
        uint64_t b = chunk_begin(t), e = b + chunk_len(t);
        if (b < e) {
            // First row whose range contains b: last r with rowptr[r] <= b.
            uint32_t r = (uint32_t)(std::upper_bound(rowptr, rowptr + rows + 1, b) - rowptr - 1);
            for (uint64_t k = b; k < e; ++k) {
                while (rowptr[r + 1] <= k) ++r;
                uint64_t pos = cur[col[k]]++;
                rowT[pos] = r;
                if (valT) valT[pos] = val[k];
            }
        }
    }
    return MLEM_OK;
}

// partitionMatrix (main.cu:593-611): segments[s] is the first row i with
// (double)rowptr[i] >= s * ((double)nnz / parts); segments[parts] = rows.
int mlem_partition(const uint64_t *rowptr, uint64_t nnz, uint32_t rows, uint32_t parts, uint32_t *segments,
                   uint32_t *segment_rows, uint32_t *segment_nnzs, uint64_t *segment_nnzs64, uint64_t *offsets) {
    if (!rowptr || !segments || !segment_rows || !segment_nnzs || !offsets)
        return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_partition: null argument");
    if (parts == 0) return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_partition: parts must be >= 1");
    segments[0] = 0;
    segments[parts] = rows;
    const double per = (double)nnz / (double)parts;
    uint32_t i = 0;
    for (uint32_t s = 0; s < parts; ++s) {
        const double target = per * (double)s;
        while (i <= rows && !((double)rowptr[i] >= target)) ++i;
        segments[s] = i;
    }
    for (uint32_t s = 0; s < parts; ++s) {
        segment_rows[s] = segments[s + 1] - segments[s];
        uint64_t n = rowptr[segments[s + 1]] - rowptr[segments[s]];
        segment_nnzs[s] = (uint32_t)n;
        if (segment_nnzs64) segment_nnzs64[s] = n;
        offsets[s] = rowptr[segments[s]];
    }
    return MLEM_OK;
}

}  // extern "C"
