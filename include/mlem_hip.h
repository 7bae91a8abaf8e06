/*
 * mlem_hip.h -- device-side C ABI of the MLEM engine (libmlem_hip.so, gfx950).
 *
 * Three layers, each replacing one interface of the reference:
 *
 *  1. Kernel-level entry points: one per __global__ of kernel.cuh:7-19, plus
 *     the fused variants (FP+ratio, BP+update) and the FP64 sensitivity.
 *     They take device pointers, 64-bit row pointers and a HIP stream
 *     (void*, NULL = the null stream), launch asynchronously and return a
 *     status.  Unlike the reference kernels (which atomicAdd into a
 *     pre-zeroed output) they OVERWRITE their outputs and, apart from
 *     mlem_bp_scatter and mlem_colsum_scatter_f64, are deterministic: the same
 *     inputs give bitwise the same outputs on every call.
 *  2. Resident matrix + reconstruction handles: the system matrix A and its
 *     transpose stay in HBM; a reconstruction owns its image/ratio buffers and
 *     one stream, so different handles may be driven from different threads.
 *  3. mlem_run(): the whole reconstruction, replacing
 *     void mlem(...) (main.cu:19,167-514) with const inputs and a status.
 *
 * Errors: every call returns an mlem_status; mlem_last_error() holds a
 * message for the calling thread.  Every HIP call inside is checked.
 */
#ifndef MLEM_HIP_H
#define MLEM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef MLEM_STATUS_DEFINED
#define MLEM_STATUS_DEFINED
typedef enum mlem_status {
    MLEM_OK = 0,
    MLEM_ERR_INVALID_ARGUMENT = 1,
    MLEM_ERR_HIP = 2,
    MLEM_ERR_OUT_OF_MEMORY = 3,
    MLEM_ERR_UNSUPPORTED = 4,
    MLEM_ERR_NO_DEVICE = 5,
    MLEM_ERR_IO = 6,
    MLEM_ERR_FORMAT = 7
} mlem_status;
#endif

const char *mlem_last_error(void);
const char *mlem_status_string(int status);
/* Library version string, e.g. "mlem-mi355x 0.1 (gfx950)". */
const char *mlem_version(void);

/* ------------------------------------------------------------------------ */
/* runtime helpers (so callers need no HIP runtime of their own)            */
/* ------------------------------------------------------------------------ */
int mlem_device_count(int *count);            /* cudaGetDeviceCount, main.cu:57-62 */
int mlem_set_device(int device);
int mlem_device_info(int device, char *name, size_t name_len,
                     uint64_t *total_mem, int *cu_count);
int mlem_device_alloc(void **ptr, size_t bytes);
int mlem_device_free(void *ptr);
int mlem_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int mlem_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int mlem_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);
int mlem_stream_create(void **stream);
int mlem_stream_destroy(void *stream);
int mlem_stream_sync(void *stream);
int mlem_device_sync(void);
/* HIP events for timing on a given stream. */
int mlem_event_create(void **event);
int mlem_event_destroy(void *event);
int mlem_event_record(void *event, void *stream);
int mlem_event_elapsed_ms(void *start, void *stop, float *ms);
/* Fill n floats with a value (a kernel, not hipMemset). */
int mlem_fill_f32(float *dst, float value, uint64_t n, void *stream);
/* dst = src over n floats with a streaming copy kernel (16-byte aligned
 * pointers): the HBM reference for bandwidth and traffic counters.  No
 * counterpart in the reference. */
int mlem_copy_f32(float *dst, const float *src, uint64_t n, void *stream);

/* ------------------------------------------------------------------------ */
/* 1. kernel-level entry points                                             */
/* ------------------------------------------------------------------------ */

/* calcCorrel (kernel.cu:11-19): fwproj[i] = (float)g[i] / fwproj[i] where
 * fwproj[i] != 0, else left as is.  In place over fwproj f32[rows]. */
int mlem_ratio(const int32_t *g, float *fwproj, uint32_t rows, void *stream);

/* calcUpdate (kernel.cu:28-39): bwproj[j] = f[j]*bwproj[j]/norm[j], or
 * f[j]*bwproj[j] where norm[j] == 0.  In place over bwproj f32[cols]. */
int mlem_update(const float *f, const float *norm, float *bwproj,
                uint32_t cols, void *stream);

/* Merge-path SpMV y = A x (calcFwProj_merge_based / calcBwProj_merge_based,
 * kernel.cu:52-105,190-286).  secsize = merge-path items per thread
 * (argv[5]/argv[6]); each 256-thread workgroup owns 256*secsize items.
 * Supported secsize: 1..32.  The workspace holds the tile partition and the
 * cross-tile carries; its size comes from mlem_merge_workspace_bytes().
 * y is f32[rows] and is overwritten (no pre-zero needed).  col and val must
 * be 16-byte aligned (hipMalloc'd arrays are), as for the CSR-vector calls. */
size_t mlem_merge_workspace_bytes(uint32_t rows, uint64_t nnz,
                                  uint32_t secsize);
int mlem_fp_merge(const uint64_t *rowptr, const uint32_t *col,
                  const float *val, const float *f, float *fwproj,
                  uint32_t secsize, uint32_t rows, uint64_t nnz,
                  void *workspace, size_t workspace_bytes, void *stream);
int mlem_bp_merge(const uint64_t *rowptrT, const uint32_t *colT,
                  const float *valT, const float *correl, float *bwproj,
                  uint32_t secsize, uint32_t cols, uint64_t nnz,
                  void *workspace, size_t workspace_bytes, void *stream);
/* mlem_fp_merge without a workspace and without allocating anything: no
 * counterpart in the reference.  For callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The partition and carries live in a fixed array of
 * MLEM_NOALLOC_SCRATCH_BYTES per device (held until the process ends), so
 * matrices needing more (mlem_merge_workspace_bytes() above it: rows + nnz
 * beyond about 1.6 million x 256 x secsize) are refused with
 * MLEM_ERR_INVALID_ARGUMENT.  Calls that overlap on one device share that
 * array and race: issue them on one stream, or wait between them.  Output
 * bitwise that of mlem_fp_merge. */
#define MLEM_NOALLOC_SCRATCH_BYTES (32u << 20)
int mlem_fp_merge_noalloc(const uint64_t *rowptr, const uint32_t *col,
                          const float *val, const float *f, float *fwproj,
                          uint32_t secsize, uint32_t rows, uint64_t nnz,
                          void *stream);

/* CSR-vector SpMV (calcFwProj_csr_vector / calcBwProj_csr_vector,
 * kernel.cu:118-153,297-336): a group of `lanes` work-items per row
 * (lanes in {4,8,16,32,64}; 0 picks from the mean row length), a
 * fixed-order shuffle reduction and a plain store.  y overwritten. */
int mlem_fp_csr_vector(const uint64_t *rowptr, const uint32_t *col,
                       const float *val, const float *f, float *fwproj,
                       uint32_t rows, uint64_t nnz, uint32_t lanes,
                       void *stream);
int mlem_bp_csr_vector(const uint64_t *rowptrT, const uint32_t *colT,
                       const float *valT, const float *correl, float *bwproj,
                       uint32_t cols, uint64_t nnz, uint32_t lanes,
                       void *stream);

/* BP without a transpose (calcBwProj_none_trans, kernel.cu:165-177,347-370):
 * bwproj = A^T correl by a float-atomic scatter over A's rows.  bwproj
 * f32[cols] is zeroed first (overwrite semantics).  Not deterministic in the
 * last bits (atomic order), like the reference. */
int mlem_bp_scatter(const uint64_t *rowptr, const uint32_t *col,
                    const float *val, const float *correl, float *bwproj,
                    uint32_t rows, uint32_t cols, uint64_t nnz, uint32_t lanes,
                    void *stream);

/* Fused FP + ratio: ratio[i] = g[i] / (A f)_i where (A f)_i != 0, else 0. */
int mlem_fp_ratio_merge(const uint64_t *rowptr, const uint32_t *col,
                        const float *val, const float *f, const int32_t *g,
                        float *ratio, uint32_t secsize, uint32_t rows,
                        uint64_t nnz, void *workspace, size_t workspace_bytes,
                        void *stream);
int mlem_fp_ratio_csr_vector(const uint64_t *rowptr, const uint32_t *col,
                             const float *val, const float *f,
                             const int32_t *g, float *ratio, uint32_t rows,
                             uint64_t nnz, uint32_t lanes, void *stream);

/* Fused BP + update on A^T: f[j] <- f[j]*bp_j/norm[j] (f[j]*bp_j where
 * norm[j] == 0), bp = A^T ratio.  f is read and written in place. */
int mlem_bp_update_merge(const uint64_t *rowptrT, const uint32_t *colT,
                         const float *valT, const float *ratio,
                         const float *norm, float *f, uint32_t secsize,
                         uint32_t cols, uint64_t nnz, void *workspace,
                         size_t workspace_bytes, void *stream);
int mlem_bp_update_csr_vector(const uint64_t *rowptrT, const uint32_t *colT,
                              const float *valT, const float *ratio,
                              const float *norm, float *f, uint32_t cols,
                              uint64_t nnz, uint32_t lanes, void *stream);

/* Sensitivity (column sums of A, calcColumnSums main.cu:558-580) in FP64 as
 * the row sums of A^T: deterministic.  sens64 f64[cols] and/or sens32
 * f32[cols] (rounded from the FP64 sum); either may be NULL. */
int mlem_colsum_f64(const uint64_t *rowptrT, const float *valT,
                    uint32_t cols, uint64_t nnz, double *sens64, float *sens32,
                    void *stream);
/* The same from A itself (no transpose) with FP64 atomics. */
int mlem_colsum_scatter_f64(const uint64_t *rowptr, const uint32_t *col,
                            const float *val, uint32_t rows, uint32_t cols,
                            uint64_t nnz, double *sens64, float *sens32,
                            void *stream);

/* ------------------------------------------------------------------------ */
/* 2. resident matrix and reconstruction handles                            */
/* ------------------------------------------------------------------------ */

typedef struct mlem_matrix mlem_matrix;
typedef struct mlem_recon mlem_recon;

/* argv[5..8] of the reference CLI (main.cu:50-55) plus tuning knobs. */
typedef struct mlem_opts {
    uint32_t spmv;        /* argv[8]: 0 merge-based, 1 CSR-vector            */
    uint32_t using_trans; /* argv[7]: 0 BP on A^T, 1 BP by scatter over A    */
    uint32_t secsize_fp;  /* argv[5]: merge-path items per thread, FP        */
    uint32_t secsize_bp;  /* argv[6]: merge-path items per thread, BP        */
    uint32_t fp_lanes;    /* CSR-vector lanes per row for FP (0 = auto)      */
    uint32_t bp_lanes;    /* CSR-vector lanes per row for BP (0 = auto)      */
    uint32_t profile;     /* 1: time FP and BP with HIP events per iteration */
    uint32_t use_graph;   /* 1: replay iterations from a captured hipGraph   */
    uint32_t bp_spmv;     /* BP algorithm: 0 = as spmv (argv[8] picks both,
                             as in the reference), 1 = merge-based,
                             2 = CSR-vector                                  */
    uint32_t tile_width;  /* merge-based kernels: 0 = blocked (segments of
                             1024 columns) with x segments staged in LDS
                             when the matrix allows it (rows sorted by
                             column), else the plain merge path;
                             1 = plain merge path; 1024 or 2048 = blocked
                             with that segment width (an error if the
                             matrix does not allow it)                       */
    uint32_t tile_rows;   /* rows per block of the blocked kernels: 0 = 2048
                             (512 when nnz < 2 x 2048 x the workgroup
                             count: a run's partial sums would cost as much
                             as its entries), or 512, 1024, 2048, 4096,
                             8192, 16384                                     */
    uint32_t tile_groups; /* workgroups per blocked SpMV (runs of equal
                             length over each row block's tiles):
                             0 = 24 per CU                                   */
    uint32_t values;      /* stored matrix values of the blocked kernels:
                             0 = f32, 1 = bf16, 2 = fp16, 3 = uf16 (an
                             unsigned 16-bit float for nonnegative matrices:
                             4 exponent and 12 mantissa bits, every value
                             from 2^-15 of its row's largest up within 2^-13
                             of itself, smaller ones stored as 2^-15 of it;
                             negative values are refused).  16-bit values
                             are value / scale rounded to nearest, with an
                             f32 scale per row of A (the row's largest
                             |value|, or 2^-14 of it for fp16); the
                             back projection uses the same codes, and the
                             sensitivity is this quantized matrix's column
                             sums in FP64 (a sensitivity passed in is not
                             used); products, sums, ratio and update stay
                             f32)                                            */
} mlem_opts;

/* Fill opts with the defaults: merge-based FP and BP (argv[8] = 0),
 * transposed BP, secsize 8, auto lanes, BP as spmv, automatic tiling, no
 * profiling, no graph. */
void mlem_opts_default(mlem_opts *opts);

/* Upload A (and A^T if rowptrT != NULL) from host arrays to `device` and keep
 * them resident.  Replaces the per-GPU cudaMalloc/cudaMemcpy of
 * main.cu:236-299 for one GPU with 64-bit row pointers. */
int mlem_matrix_create(int device, uint32_t rows, uint32_t cols, uint64_t nnz,
                       const uint64_t *rowptr, const uint32_t *col,
                       const float *val, const uint64_t *rowptrT,
                       const uint32_t *colT, const float *valT,
                       mlem_matrix **out);
/* Build A^T on the device from the resident A (stable counting sort,
 * bit-exact with the host transpose) instead of uploading one.  Replaces
 * sptrans_scanTrans_specialized (sptrans.h:397-545) as called at
 * main.cu:99-105.  The one sort needs 20 bytes of scratch per nonzero; when
 * that is not free (the 100 GB matrices) the sort runs over column ranges
 * within the free memory, with the same result.  Refused
 * (MLEM_ERR_INVALID_ARGUMENT) while a reconstruction or batch of the matrix
 * exists: they hold the transpose's arrays, tiles and captured graphs. */
int mlem_matrix_build_transpose(mlem_matrix *m);
/* The same within scratch_bytes of scratch (0 = what is free); *ranges (may
 * be NULL) = column ranges used, 0 for the single sort.  MLEM_ERR_OUT_OF_MEMORY
 * when one column's nonzeros do not fit. */
int mlem_matrix_build_transpose_limited(mlem_matrix *m, uint64_t scratch_bytes, uint32_t *ranges);
/* A row order for the batched layouts: order[i] (a permutation of
 * 0..rows-1) is the place of row i (a LOR) of A.  Batches created afterwards
 * (mlem_batch_create, the server's) build their tiles from A with its rows
 * in that order and from that matrix's transpose, and place each sinogram
 * in it; images are the same up to FP32 rounding (sums in another order).
 * LORs that cross a voxel then fall into fewer column segments of A^T's
 * tiles, whose rows are padded to whole quads of entries: on the 3D scanner
 * matrix the stored entries of A^T drop from +31.6% to about +5% of its
 * nonzeros with the order of mlem_lor_order (include/mlem_host.h).  NULL
 * clears it.  Single reconstructions do not use it.  Refused while
 * reconstructions or batches use the matrix, and on tiles-only matrices.
 * No counterpart in the reference (its kernels take A's rows as stored). */
int mlem_matrix_set_row_order(mlem_matrix *m, const uint32_t *order);
/* A matrix resident as tiles only: the blocked layout of A and of A^T
 * (`values` as in mlem_opts; per_pass 0 or 1: the tiles of single
 * reconstructions, tile width 0 = 1024 or 2048, tile_rows 0 as in mlem_opts;
 * per_pass 4 or 8: those of batches with that many per pass, widths and
 * rows as in mlem_batch_opts) and the FP64 sensitivity (of the quantized matrix for
 * 16-bit values), without the CSR copies, for matrices whose
 * two CSR copies leave no room for the tiles (the 100 GB matrix: 265 of
 * 288 GB).  Replaces the upload of main.cu:236-299 plus the transpose of
 * main.cu:99-105 for that case.  A is uploaded, A^T's tiles are built from
 * column ranges of A (whole row blocks of A^T) and the sensitivity from their
 * row sums; A's copy is freed; A's tiles are built from row chunks of the
 * host CSR (read a second time).  The tiles are byte for byte those of the
 * resident path.  chunk_bytes bounds the scratch of each range / chunk
 * (0 = what is free).  A reconstruction on it (per_pass 1) takes merge-based
 * FP and BP on A^T (spmv 0, using_trans 0) with the matrix's tile shape and
 * values, a batch (per_pass 4 or 8) the matrix's per_pass, shape and values;
 * the plain kernels, batches, the autotuner and the transpose calls refuse
 * it.  MLEM_ERR_OUT_OF_MEMORY when the tiles or one row block's scratch do
 * not fit. */
int mlem_matrix_create_tiled(int device, uint32_t rows, uint32_t cols, uint64_t nnz,
                             const uint64_t *rowptr, const uint32_t *col,
                             const float *val, uint32_t tile_width,
                             uint32_t tile_rows, uint32_t values,
                             uint32_t per_pass, uint64_t chunk_bytes,
                             mlem_matrix **out);

/* Load A straight from a csr4 file into a resident matrix (Csr4Matrix +
 * mapRows + csr_format_for_cuda + the H2D copies of main.cu:71-79,263-272,
 * without a host CSR): the mapped AoS data section is copied by `threads`
 * host threads (<= 0: all cores) into pinned staging buffers, streamed to
 * the device and split there into col u32[nnz] / val f32[nnz]; copies,
 * transfers and the split overlap.  The row pointers are the file's
 * inclusive row ends with a leading 0.  A column index >= cols is an
 * MLEM_ERR_FORMAT error.  stats may be NULL. */
typedef struct mlem_load_stats {
    uint64_t bytes;  /* data-section bytes streamed to the device */
    double open_s;   /* open, validate and map the file */
    double stream_s; /* copy + H2D + split of the data section */
    double total_s;
} mlem_load_stats;
int mlem_matrix_load_csr4(int device, const char *path, int threads,
                          mlem_matrix **out, mlem_load_stats *stats);
/* Refused (MLEM_ERR_INVALID_ARGUMENT, nothing freed) while a reconstruction
 * or batch of the matrix exists: destroy those first. */
int mlem_matrix_destroy(mlem_matrix *m);
int mlem_matrix_info(const mlem_matrix *m, uint32_t *rows, uint32_t *cols,
                     uint64_t *nnz, int *has_transpose, uint64_t *device_bytes);
/* Device pointers of the resident arrays (any may be NULL). */
int mlem_matrix_device_arrays(const mlem_matrix *m, const uint64_t **rowptr,
                              const uint32_t **col, const float **val,
                              const uint64_t **rowptrT, const uint32_t **colT,
                              const float **valT);
/* Download A^T (built or uploaded) to host arrays. */
int mlem_matrix_download_transpose(const mlem_matrix *m, uint64_t *colptrT,
                                   uint32_t *rowT, float *valT);
/* FP64 sensitivity on the GPU; sens32 f32[cols] and/or sens64 f64[cols]
 * host outputs (either may be NULL). */
int mlem_matrix_sensitivity(const mlem_matrix *m, float *sens32,
                            double *sens64);

/* A reconstruction over a resident matrix: g int32[rows] counts, sens
 * f32[cols] (NULL: FP64 sensitivity on the GPU), f0 f32[cols] start image. */
int mlem_recon_create(const mlem_matrix *m, const int32_t *g,
                      const float *sens, const float *f0,
                      const mlem_opts *opts, mlem_recon **out);
/* Enqueue `iters` MLEM iterations (main.cu:329-449) on the handle's stream. */
int mlem_recon_iterate(mlem_recon *r, uint32_t iters);
int mlem_recon_sync(mlem_recon *r);
int mlem_recon_get_image(mlem_recon *r, float *f_out);
/* Replace the image (f32[cols]) to restart from another x. */
int mlem_recon_set_image(mlem_recon *r, const float *f);
/* Sum over iterations of the FP and BP kernel times (HIP events on the
 * handle's stream; needs opts.profile) and the number of iterations timed.
 * The FP time covers FP+ratio (+ the merge fix-up), the BP time covers
 * BP+update (+ fix-up, or scatter+update). */
int mlem_recon_kernel_times(mlem_recon *r, double *fp_ms, double *bp_ms,
                            uint32_t *iters_timed);
/* The same split in four: ms[0] the FP SpMV kernel (the blocked kernel's
 * partial sums; merge-based or CSR-vector FP with its fused ratio), ms[1]
 * what follows it in FP (the blocked path's reduction + ratio, else ~0),
 * ms[2] the BP SpMV (or scatter) kernel, ms[3] the rest of BP (reduction +
 * update, or the scatter path's update). */
int mlem_recon_phase_times(mlem_recon *r, double ms[4], uint32_t *iters_timed);
int mlem_recon_reset_times(mlem_recon *r);
/* Poisson log-likelihood sum_i g_i log(Ax)_i - (Ax)_i of the current image,
 * computed on the GPU in FP64 (rows with (Ax)_i <= 0 contribute -(Ax)_i). */
int mlem_recon_loglik(mlem_recon *r, double *ll);
/* The handle's stream (a hipStream_t). */
void *mlem_recon_stream(mlem_recon *r);
int mlem_recon_destroy(mlem_recon *r);

/* ------------------------------------------------------------------------ */
/* 2b. batched reconstructions (many sinograms, one resident matrix)        */
/* ------------------------------------------------------------------------ */

/* `slots` reconstructions over one matrix (A and A^T resident) iterated
 * together: every pass of the blocked FP / BP kernels over the matrix serves
 * `per_pass` slots at once, so the matrix is read once per per_pass
 * reconstructions.  Each slot's arithmetic does not depend on the other
 * slots' data: its image is the same whatever else is in the batch (with the
 * same options).  The update skips inactive slots, whose images stay as
 * they are.  No counterpart in the reference (one sinogram per process). */
typedef struct mlem_batch mlem_batch;
typedef struct mlem_batch_opts {
    uint32_t values;      /* stored matrix values: 0 f32, 1 bf16, 2 fp16,
                             3 uf16 (as in mlem_opts)                     */
    uint32_t per_pass;    /* slots per pass over the matrix: 1, 4, 8
                             (0 = 4)                                       */
    uint32_t tile_width;  /* 0 = automatic (512 for 8 per pass, else 1024) */
    uint32_t tile_rows;   /* 0 = automatic (2048 for 1 per pass, else 1024;
                             512 for small matrices, as in mlem_opts)        */
    uint32_t tile_groups; /* workgroups per blocked SpMV, 0 = automatic    */
    uint32_t tile_threads; /* work-items per workgroup of the kernel for
                              4 or 8 per pass: 128, 256 or 512 (each takes
                              512 / tile_threads groups of 8 entries of a
                              window), 0 = automatic (512)               */
    uint32_t use_graph;    /* 1: each mlem_batch_iterate call replays one
                              iteration captured as a hipGraph (captured
                              again when the set of passes holding an
                              active slot changes)                        */
} mlem_batch_opts;
void mlem_batch_opts_default(mlem_batch_opts *opts);
/* The sensitivity (shared by all slots) is the matrix's FP64 column sums
 * (with 16-bit values: those of the quantized matrix). */
int mlem_batch_create(mlem_matrix *m, uint32_t slots,
                      const mlem_batch_opts *opts, mlem_batch **out);
/* Load a reconstruction into a slot (g int32[rows], f0 f32[cols]) and make
 * it active. */
int mlem_batch_set(mlem_batch *b, uint32_t slot, const int32_t *g,
                   const float *f0);
/* Stop updating a slot (its image stays readable). */
int mlem_batch_deactivate(mlem_batch *b, uint32_t slot);
/* Enqueue `iters` iterations of every pass that holds an active slot. */
int mlem_batch_iterate(mlem_batch *b, uint32_t iters);
int mlem_batch_sync(mlem_batch *b);
int mlem_batch_get_image(mlem_batch *b, uint32_t slot, float *f_out);
int mlem_batch_info(const mlem_batch *b, uint32_t *slots, uint32_t *per_pass,
                    uint32_t *passes, uint64_t *device_bytes);
void *mlem_batch_stream(mlem_batch *b);
int mlem_batch_destroy(mlem_batch *b);

/* ------------------------------------------------------------------------ */
/* 2b'. autotuner                                                           */
/* ------------------------------------------------------------------------ */

/* Times the blocked FP and BP kernels of the matrix over a few tile shapes
 * (width x rows) and workgroup counts for `per_pass` right-hand sides (1, 4
 * or 8) and `values` (0 f32, 1 bf16, 2 fp16, 3 uf16), and reports the fastest for
 * mlem_opts / mlem_batch_opts (tile_width, tile_rows, tile_groups).  Tiles
 * built for candidates that were not chosen are freed again.  No counterpart
 * in the reference (its section sizes come from argv[5]/[6]). */
typedef struct mlem_tune_result {
    uint32_t tile_width, tile_rows, tile_groups;
    uint32_t tile_threads; /* for mlem_batch_opts (0 for 1 per pass)       */
    uint32_t candidates;  /* shape x workgroup-count pairs timed            */
    double fp_ms, bp_ms;  /* per launch of the chosen one, partials + reduce */
} mlem_tune_result;
int mlem_autotune(mlem_matrix *m, uint32_t per_pass, uint32_t values,
                  mlem_tune_result *out);

/* ------------------------------------------------------------------------ */
/* 2c. serving: continuous batching over resident matrices                  */
/* ------------------------------------------------------------------------ */

/* A request loop over several resident matrices, each with a batch of slots
 * (mlem_batch_*).  Admission: a matrix and its batch are registered only if
 * they fit in what is left of the HBM budget (0 = no limit); requests queue
 * per matrix and take a free slot of their batch between two of its
 * iterations.  mlem_server_step starts one iteration on every batch that has
 * active slots and none in flight (each batch on its own stream: batches of
 * different matrices overlap and go at their own pace), waits up to
 * max_wait_us for one to end and retires the requests that reached their
 * iteration count.  A request's image is
 * the same as when it runs alone (same batch options).  Driven by the
 * caller's thread; not thread-safe.  No counterpart in the reference (one
 * reconstruction per process). */
typedef struct mlem_server mlem_server;
int mlem_server_create(uint64_t hbm_budget_bytes,
                       const mlem_batch_opts *batch_opts, mlem_server **out);
/* The matrix (A and A^T resident) stays owned by the caller. */
int mlem_server_add_matrix(mlem_server *s, mlem_matrix *m, uint32_t max_slots,
                           uint32_t *matrix_id);
/* g int32[rows] and f0 f32[cols] are copied. */
int mlem_server_submit(mlem_server *s, uint32_t matrix_id, const int32_t *g,
                       const float *f0, uint32_t iters, uint64_t *request_id);
/* *active = requests in slots after the step, *queued = waiting (both may be
 * NULL).  max_wait_us = 0 only collects iterations that already ended. */
int mlem_server_step(mlem_server *s, uint32_t max_wait_us, uint32_t *active, uint32_t *queued);
/* *done = 1 with the image (f32[cols], may be NULL) and the latency from
 * submission when finished (the request is then released), else 0. */
int mlem_server_poll(mlem_server *s, uint64_t request_id, float *f_out,
                     double *latency_s, int *done);
int mlem_server_info(const mlem_server *s, uint64_t *budget, uint64_t *used,
                     uint32_t *matrices, uint64_t *steps);
/* Work done so far: *pass_sweeps = passes over a matrix started (a pass
 * costs a full sweep of its matrix however few of its slots are active),
 * *slot_moves = requests moved to a lower pass so that a pass empties (the
 * requests of the highest pass with active slots move whenever they fit in
 * free slots of lower passes; the images stay bitwise the same).  Either may
 * be NULL.  No counterpart in the reference. */
int mlem_server_stats(const mlem_server *s, uint64_t *pass_sweeps, uint64_t *slot_moves);
int mlem_server_destroy(mlem_server *s);

/* ------------------------------------------------------------------------ */
/* 3. whole reconstruction                                                  */
/* ------------------------------------------------------------------------ */

/* mlem() of main.cu:167-514 on one GPU: iterate `iters` times from f0 and
 * write the image to f_out (host arrays; the caller keeps ownership, nothing
 * is mutated).  sens NULL: FP64 sensitivity on the GPU. */
int mlem_run(const mlem_matrix *m, const int32_t *g, const float *sens,
             const float *f0, float *f_out, uint32_t iters,
             const mlem_opts *opts);

#ifdef __cplusplus
}
#endif
#endif /* MLEM_HIP_H */
