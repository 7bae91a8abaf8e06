/*
 * mlem_oracle.c -- CPU restatement of the reference MLEM path.
 *
 * TEST INFRASTRUCTURE ONLY.  Results are checked against it: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * liboracle.so, and never as the thing measured or shipped.  The engine's
 * product path (libmlem_host.so, libmlem_hip.so) does not link or call it.
 *
 * Every function restates the reference code it cites (paths relative to the
 * reference repository).  How each one is pinned (what a test checks it
 * against) is listed in oracle/README.md:
 *   oracle_csr_from_csr4      main.cu:518-554           csr_format_for_cuda
 *   oracle_colsum_f64/_f32    main.cu:558-580           calcColumnSums
 *   oracle_initial_image      main.cu:83-97             init = sum_g / sum_norm
 *   oracle_sptrans            sptrans.h:195-243,397-545 ScanTrans (specialized)
 *   oracle_partition          main.cu:593-611           partitionMatrix
 *   oracle_merge_coord        kernel.cu:202-229         merge_based_start search
 *   oracle_merge_spmv_f64     kernel.cu:190-286         merge-based SpMV walk
 *   oracle_spmv_f64           kernel.cu:297-336         CSR SpMV (row sums)
 *   oracle_mlem_f64           main.cu:329-449, kernel.cu:11-39, 347-370
 *   oracle_mlem_f64_omp       the same, B images, OpenMP, bitwise equal to it
 *   oracle_loglik_f64         Poisson log-likelihood (the MLEM objective)
 *   oracle_mlem_f32_omp       main.cu:329-449 as an FP32 OpenMP CPU loop
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* csr_format_for_cuda (main.cu:518-554): rowptr[0] = 0, rowptr[r+1] =
 * rowptr[r] + elementsInRow(r) (serial, :526-530), then each row's AoS
 * {u32 col, f32 val} entries copied in order (:532-553).  row_ends are the
 * csr4 inclusive row ends (csr4matrix.cpp:267-271).  64-bit counters are used
 * where the reference's `unsigned int index` wraps above 2^32 nnz. */
void oracle_csr_from_csr4(uint32_t rows, const uint64_t *row_ends, const uint8_t *aos, uint64_t *rowptr,
                          uint32_t *col, float *val) {
    rowptr[0] = 0;
    for (uint32_t r = 0; r < rows; ++r) {
        uint64_t in_row = (r == 0) ? row_ends[0] : row_ends[r] - row_ends[r - 1];
        rowptr[r + 1] = rowptr[r] + in_row;
    }
    for (uint32_t r = 0; r < rows; ++r) {
        uint64_t idx = 0;
        uint64_t begin = (r == 0) ? 0 : row_ends[r - 1]; /* beginRow2, csr4matrix.hpp:138-141 */
        for (uint64_t k = begin; k < row_ends[r]; ++k, ++idx) {
            memcpy(&col[rowptr[r] + idx], aos + 8 * k, 4);
            memcpy(&val[rowptr[r] + idx], aos + 8 * k + 4, 4);
        }
    }
}

/* calcColumnSums (main.cu:558-580) in FP64: sens[c] = sum of column c. */
void oracle_colsum_f64(const uint64_t *rowptr, const uint32_t *col, const float *val, uint32_t rows, uint32_t cols,
                       double *sens) {
    for (uint32_t c = 0; c < cols; ++c) sens[c] = 0.0;
    for (uint32_t r = 0; r < rows; ++r)
        for (uint64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) sens[col[k]] += (double)val[k];
}

/* calcColumnSums in FP32 with one OpenMP thread: res[col] += value in row order. */
void oracle_colsum_f32(const uint64_t *rowptr, const uint32_t *col, const float *val, uint32_t rows, uint32_t cols,
                       float *norm) {
    for (uint32_t c = 0; c < cols; ++c) norm[c] = 0.0f;
    for (uint32_t r = 0; r < rows; ++r)
        for (uint64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) norm[col[k]] += val[k];
}

/* main.cu:83-97: float sum_norm (serial), int sum_g, init = sum_g / sum_norm. */
float oracle_initial_image(const int32_t *g, uint32_t rows, const float *norm, uint32_t cols) {
    float sum_norm = 0.0f;
    for (uint32_t i = 0; i < cols; ++i) sum_norm += norm[i];
    uint32_t sum_g = 0; /* int arithmetic of the reference, wrap made explicit */
    for (uint32_t i = 0; i < rows; ++i) sum_g += (uint32_t)g[i];
    return (float)(int32_t)sum_g / sum_norm;
}

/* ---- ScanTrans (sptrans.h) ------------------------------------------------ */

#define SCAN_BSIZE 2048 /* _SCAN_BSIZE, sptrans.h:25 */

/* scan_kernel_specialized without AVX2 (sptrans.h:195-214). */
static void scan_kernel(uint64_t *in, uint64_t *out, uint64_t N, int inclusive, uint64_t base) {
    if (N == 0) return;
    if (inclusive) {
        out[0] = in[0] + base;
        for (uint64_t i = 1; i < N; i++) out[i] = out[i - 1] + in[i];
    } else {
        uint64_t sum = base;
        for (uint64_t i = 0; i < N; i++) {
            uint64_t tmp = in[i];
            out[i] = sum;
            sum += tmp;
        }
    }
}

/* scan_omp_specialzed (sptrans.h:217-243): blocks of 2048, block sums taken
 * with std::accumulate(in+st, in+ed, 0) -- an `int` accumulator (wraps
 * mod 2^32 and is sign-extended into the u64 sums[]) -- an exclusive scan of
 * the block sums, then each block scanned from its base. */
static void scan_omp(uint64_t *in, uint64_t *out, uint64_t N, int inclusive, uint64_t base) {
    uint64_t nblocks = N ? 1 + (N - 1) / SCAN_BSIZE : 0; /* NBLOCKS, sptrans.h:24 */
    if (nblocks <= 2) {
        scan_kernel(in, out, N, inclusive, base);
        return;
    }
    uint64_t *sums = (uint64_t *)malloc(nblocks * sizeof(uint64_t));
    for (uint64_t i = 0; i < nblocks; i++) {
        uint64_t st = i * SCAN_BSIZE, ed = st + SCAN_BSIZE < N ? st + SCAN_BSIZE : N;
        int32_t sum = 0;
        for (uint64_t k = st; k < ed; ++k) sum = (int32_t)(uint32_t)((uint64_t)(int64_t)sum + in[k]);
        sums[i] = (uint64_t)(int64_t)sum;
    }
    scan_omp(sums, sums, nblocks, 0, base);
    for (uint64_t i = 0; i < nblocks; i++) {
        uint64_t st = i * SCAN_BSIZE, ed = st + SCAN_BSIZE < N ? st + SCAN_BSIZE : N;
        scan_kernel(in + st, out + st, ed - st, inclusive, sums[i]);
    }
    free(sums);
}

/* sptrans_scanTrans_specialized (sptrans.h:397-545) with `procs` threads run
 * one after another.  cscColPtr must be zeroed by the caller (main.cu:101).
 * cscVal may be NULL (sptrans.h:492,517). */
int oracle_sptrans(uint32_t m, uint32_t n, uint64_t nnz, const uint64_t *csrRowPtr, const uint32_t *csrColIdx,
                   const float *csrVal, uint32_t *cscRowIdx, uint64_t *cscColPtr, float *cscVal, uint32_t procs) {
    if (procs == 0) return -1;
    uint64_t size = nnz / procs, left = nnz % procs;
    uint64_t *csrRowUnroll = (uint64_t *)malloc((nnz ? nnz : 1) * sizeof(uint64_t));
    uint64_t *intra = (uint64_t *)malloc((nnz ? nnz : 1) * sizeof(uint64_t));
    uint64_t *wb_index = (uint64_t *)malloc((nnz ? nnz : 1) * sizeof(uint64_t));
    uint64_t *inter = (uint64_t *)calloc(((uint64_t)procs + 1) * n + 1, sizeof(uint64_t));
    if (!csrRowUnroll || !intra || !wb_index || !inter) {
        free(csrRowUnroll);
        free(intra);
        free(wb_index);
        free(inter);
        return -2;
    }
    /* chunk of thread tid (sptrans.h:425-435) */
#define CHUNK(tid, start, len)                                  \
    do {                                                        \
        start = size * (tid);                                   \
        len = size;                                             \
        if (left != 0) {                                        \
            if ((tid) < left) {                                 \
                len += 1;                                       \
                start = (tid) * size + (tid);                   \
            } else {                                            \
                start = (tid) * size + left;                    \
            }                                                   \
        }                                                       \
    } while (0)
    /* P1 (sptrans.h:420-442): intra[k] = inter[t+1][col]++ */
    for (uint64_t tid = 0; tid < procs; ++tid) {
        uint64_t start, len;
        CHUNK(tid, start, len);
        uint64_t inter_start = n + tid * n;
        for (uint64_t i = 0; i < len; i++) {
            uint64_t index = inter_start + csrColIdx[start + i];
            intra[start + i] = inter[index];
            inter[index]++;
        }
    }
    /* P2 (sptrans.h:444-449): prefix over threads */
    for (uint64_t i = 0; i < n; i++)
        for (uint64_t j = 2; j < (uint64_t)procs + 1; j++) inter[i + n * j] += inter[i + n * (j - 1)];
    /* P3 (sptrans.h:451-457) */
    for (uint64_t i = 0; i < n; i++) cscColPtr[i + 1] = inter[n * procs + i];
    scan_omp(cscColPtr, cscColPtr, (uint64_t)n + 1, 1, 0);
    /* P4 (sptrans.h:459-466): row of every nonzero */
    for (uint64_t i = 0; i < m; i++)
        for (uint64_t j = 0; j < csrRowPtr[i + 1] - csrRowPtr[i]; j++) csrRowUnroll[csrRowPtr[i] + j] = i;
    /* P5 (sptrans.h:468-490) */
    for (uint64_t tid = 0; tid < procs; ++tid) {
        uint64_t start, len;
        CHUNK(tid, start, len);
        uint64_t inter_start = tid * n;
        for (uint64_t i = 0; i < len; i++) {
            uint64_t colIdx = csrColIdx[start + i];
            wb_index[start + i] = cscColPtr[colIdx] + inter[inter_start + colIdx] + intra[start + i];
        }
    }
    /* P6 (sptrans.h:492-540) */
    for (uint64_t tid = 0; tid < procs; ++tid) {
        uint64_t start, len;
        CHUNK(tid, start, len);
        for (uint64_t i = 0; i < len; i++) {
            uint64_t off = wb_index[start + i];
            if (cscVal) cscVal[off] = csrVal[start + i];
            cscRowIdx[off] = (uint32_t)csrRowUnroll[start + i];
        }
    }
#undef CHUNK
    free(csrRowUnroll);
    free(inter);
    free(intra);
    free(wb_index);
    return 0;
}

/* partitionMatrix (main.cu:593-611). */
void oracle_partition(const uint64_t *rowptr, uint64_t nnz, uint32_t rows, uint32_t D, uint32_t *segments,
                      uint32_t *segment_rows, uint32_t *segment_nnzs, uint64_t *offsets) {
    segments[0] = 0;
    segments[D] = rows;
    uint32_t i = 0;
    double nnzs_per_segment = ((double)nnz / (double)D);
    for (uint32_t segment = 0; segment < D; segment++) {
        for (; i <= rows; i++)
            if ((double)rowptr[i] >= nnzs_per_segment * segment) break;
        segments[segment] = i;
    }
    for (uint32_t segment = 0; segment < D; segment++) {
        segment_rows[segment] = segments[segment + 1] - segments[segment];
        segment_nnzs[segment] = (uint32_t)(rowptr[segments[segment + 1]] - rowptr[segments[segment]]);
        offsets[segment] = rowptr[segments[segment]];
    }
}

/* merge_based_start's search (kernel.cu:202-229) for diagonal `start`, with
 * 64-bit indices.  The probe at i == rows (which reads rowptr[rows+1] there)
 * ends the loop at i == rows whatever it reads, so it is not evaluated. */
void oracle_merge_coord(const uint64_t *rowptr, uint32_t rows, int64_t start, int64_t *out_i, int64_t *out_j) {
    int64_t lefti = 0, righti = rows;
    int64_t nexti = righti / 2;
    int64_t nextj = start - nexti;
    int64_t i = 0, j = start;
    while (i != nexti) {
        i = nexti;
        j = nextj;
        if (i == (int64_t)rows)
            righti = i; /* any outcome leaves nexti == rows */
        else if ((int64_t)rowptr[i + 1] > j - 1)
            righti = i;
        else
            lefti = i + 1;
        nexti = (lefti + righti) / 2;
        nextj = start - nexti;
    }
    *out_i = i;
    *out_j = j;
}

/* calcFwProj_merge_based (kernel.cu:52-72,190-286) in FP64: every thread
 * index * secSize starts a walk of secSize merge items whose row partials
 * are added into y (zeroed first).  Threads: ceil((rows+nnz)/secSize). */
void oracle_merge_spmv_f64(const uint64_t *rowptr, const uint32_t *col, const float *val, const float *x,
                           uint32_t rows, uint64_t nnz, uint32_t secSize, double *y) {
    for (uint32_t r = 0; r < rows; ++r) y[r] = 0.0;
    int64_t total = (int64_t)rows + (int64_t)nnz;
    for (int64_t start = 0; start < total; start += secSize) {
        int64_t i, j;
        oracle_merge_coord(rowptr, rows, start, &i, &j);
        int64_t end = i + j + secSize;
        if (end > total) end = total;
        double acc = 0.0;
        while (i + j < end) {
            if ((int64_t)rowptr[i + 1] > j) {
                acc += (double)val[j] * (double)x[col[j]];
                j++;
            } else {
                y[i] += acc;
                i++;
                acc = 0.0;
            }
        }
        if (acc != 0.0) y[i] += acc;
    }
}

/* y = A x in FP64 (the row sums mat_vec_mul_csr_vector forms, kernel.cu:297-336). */
void oracle_spmv_f64(const uint64_t *rowptr, const uint32_t *col, const float *val, const double *x, uint32_t rows,
                     double *y) {
    for (uint32_t r = 0; r < rows; ++r) {
        double s = 0.0;
        for (uint64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) s += (double)val[k] * x[col[k]];
        y[r] = s;
    }
}

/* The MLEM loop of main.cu:329-449 in FP64 over A only:
 *   fp = A f                                     (calcFwProj_*)
 *   ratio_i = g_i / fp_i if fp_i != 0, else fp_i (calcCorrel, kernel.cu:11-19)
 *   bp = A^T ratio                               (calcBwProj_* / none_trans)
 *   f_j = f_j bp_j / sens_j, or f_j bp_j if sens_j == 0  (calcUpdate, kernel.cu:28-39)
 * f holds f0 on entry and the image after `iters` iterations on exit. */
void oracle_mlem_f64(const uint64_t *rowptr, const uint32_t *col, const float *val, uint32_t rows, uint32_t cols,
                     const int32_t *g, const double *sens, double *f, uint32_t iters) {
    double *fp = (double *)malloc(((size_t)rows + 1) * sizeof(double));
    double *bp = (double *)malloc(((size_t)cols + 1) * sizeof(double));
    for (uint32_t it = 0; it < iters; ++it) {
        oracle_spmv_f64(rowptr, col, val, f, rows, fp);
        for (uint32_t i = 0; i < rows; ++i)
            if (fp[i] != 0.0) fp[i] = (double)g[i] / fp[i];
        for (uint32_t c = 0; c < cols; ++c) bp[c] = 0.0;
        for (uint32_t r = 0; r < rows; ++r)
            for (uint64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) bp[col[k]] += (double)val[k] * fp[r];
        for (uint32_t c = 0; c < cols; ++c) f[c] = (sens[c] == 0.0) ? f[c] * bp[c] : f[c] * bp[c] / sens[c];
    }
    free(fp);
    free(bp);
}

/* oracle_mlem_f64 for B images at once, for matrices of 1e9+ entries: FP
 * parallel over rows of A, BP parallel over rows of A^T (the stable
 * transpose, so row c of A^T lists column c's entries in A's row order).
 * Every sum is formed by one thread in the serial oracle's order -- FP in
 * k order along the row, BP in increasing row of A, which is the order in
 * which oracle_mlem_f64's scatter adds into bp[c] -- so each image is
 * bitwise the one oracle_mlem_f64 gives for it, whatever the thread count.
 * Vectors are interleaved, image b of voxel c at f[c * B + b], counts of
 * image b of row r at g[r * B + b]; B <= 64. */
int oracle_mlem_f64_omp(const uint64_t *rowptr, const uint32_t *col, const float *val, const uint64_t *rowptrT,
                        const uint32_t *colT, const float *valT, uint32_t rows, uint32_t cols, uint32_t B,
                        const int32_t *g, const double *sens, double *f, uint32_t iters, int threads) {
    if (B == 0 || B > 64) return -1;
    double *ratio = (double *)malloc(((size_t)rows * B + 1) * sizeof(double));
    if (!ratio) return -2;
    if (threads > 0) omp_set_num_threads(threads);
    for (uint32_t it = 0; it < iters; ++it) {
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t r = 0; r < (int64_t)rows; ++r) {
            double s[64];
            for (uint32_t b = 0; b < B; ++b) s[b] = 0.0;
            for (uint64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
                const double v = (double)val[k];
                const double *fx = f + (size_t)col[k] * B;
                for (uint32_t b = 0; b < B; ++b) s[b] += v * fx[b];
            }
            for (uint32_t b = 0; b < B; ++b)
                ratio[(size_t)r * B + b] = (s[b] != 0.0) ? (double)g[(size_t)r * B + b] / s[b] : s[b];
        }
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t c = 0; c < (int64_t)cols; ++c) {
            double s[64];
            for (uint32_t b = 0; b < B; ++b) s[b] = 0.0;
            for (uint64_t k = rowptrT[c]; k < rowptrT[c + 1]; ++k) {
                const double v = (double)valT[k];
                const double *rx = ratio + (size_t)colT[k] * B;
                for (uint32_t b = 0; b < B; ++b) s[b] += v * rx[b];
            }
            double *fc = f + (size_t)c * B;
            for (uint32_t b = 0; b < B; ++b) fc[b] = (sens[c] == 0.0) ? fc[b] * s[b] : fc[b] * s[b] / sens[c];
        }
    }
    free(ratio);
    return 0;
}

/* Poisson log-likelihood sum_i g_i log(y_i) - y_i, y = A f (terms with
 * y_i <= 0 contribute -y_i). */
double oracle_loglik_f64(const uint64_t *rowptr, const uint32_t *col, const float *val, uint32_t rows,
                         const int32_t *g, const double *f) {
    double s = 0.0;
    for (uint32_t r = 0; r < rows; ++r) {
        double y = 0.0;
        for (uint64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) y += (double)val[k] * f[col[k]];
        s += (y > 0.0) ? (double)g[r] * log(y) - y : -y;
    }
    return s;
}

/* The same loop in FP32 with OpenMP over rows of A (FP) and rows of A^T
 * (BP): the CPU baseline, same arrays as the GPU path.  Returns seconds of
 * wall time spent in the `iters` iterations. */
double oracle_mlem_f32_omp(const uint64_t *rowptr, const uint32_t *col, const float *val, const uint64_t *rowptrT,
                           const uint32_t *colT, const float *valT, uint32_t rows, uint32_t cols, const int32_t *g,
                           const float *sens, float *f, uint32_t iters, int threads) {
    float *ratio = (float *)malloc(((size_t)rows + 1) * sizeof(float));
    if (threads > 0) omp_set_num_threads(threads);
    double t0 = omp_get_wtime();
    for (uint32_t it = 0; it < iters; ++it) {
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t r = 0; r < (int64_t)rows; ++r) {
            float s = 0.0f;
            for (uint64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) s += val[k] * f[col[k]];
            ratio[r] = (s != 0.0f) ? (float)g[r] / s : s;
        }
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t c = 0; c < (int64_t)cols; ++c) {
            float s = 0.0f;
            for (uint64_t k = rowptrT[c]; k < rowptrT[c + 1]; ++k) s += valT[k] * ratio[colT[k]];
            float v = f[c] * s;
            f[c] = (sens[c] == 0.0f) ? v : v / sens[c];
        }
    }
    double t1 = omp_get_wtime();
    free(ratio);
    return t1 - t0;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }
