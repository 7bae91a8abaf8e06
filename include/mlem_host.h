/*
 * mlem_host.h -- host-side C ABI of the MLEM engine (libmlem_host.so).
 *
 * Everything here runs on the CPU and needs no GPU: the csr4 system-matrix
 * reader/writer, the raw int32/float32 vector files, the AoS->SoA CSR build,
 * the f32 column sums, the stable counting-sort transpose (bit-exact with
 * sptrans' ScanTrans), the nnz-balanced partition, the x0 initialisation and
 * the synthetic scanner/phantom/sinogram generator.
 *
 * Conventions: every function returns an mlem_status (0 = MLEM_OK); on error
 * mlem_host_last_error() returns a message for the calling thread.  Arrays are
 * plain host pointers owned by the caller unless the function says otherwise.
 * Row pointers are uint64 (rowptr[0] = 0, rowptr[rows] = nnz), columns uint32,
 * values float32 -- the layout main.cu:76-78 hands to mlem().
 */
#ifndef MLEM_HOST_H
#define MLEM_HOST_H

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

/* Message describing the last failed call on this thread ("" if none). */
const char *mlem_host_last_error(void);
const char *mlem_host_status_string(int status);

/* ------------------------------------------------------------------------ */
/* csr4 system-matrix file (helper_files_common/csr4matrix.{hpp,cpp})        */
/* ------------------------------------------------------------------------ */

/* Scanner geometry block carried in the csr4 header
 * (csr4matrix.cpp:109-139; scannerconfig.hpp:55-103).  Variable-length parts
 * are capped: layers <= 16, block gaps <= 255, ring gaps <= 255. */
#define MLEM_MAX_LAYERS 16
#define MLEM_MAX_GAPS 255
typedef struct mlem_scanner {
    uint32_t n_layers;
    float layer[MLEM_MAX_LAYERS][5]; /* width, height, depth, radius, mu */
    uint32_t n_blocks;
    uint32_t blocksize;              /* blocksize-1 gaps follow */
    float block_gap[MLEM_MAX_GAPS];
    uint32_t n_rings;                /* n_rings-1 gaps follow */
    float ring_gap[MLEM_MAX_GAPS];
    float voxel_size[3];             /* x, y, z */
    uint32_t grid[3];                /* nx, ny, nz */
} mlem_scanner;

typedef struct mlem_csr4 mlem_csr4; /* an open, memory-mapped csr4 file */

/* Open and validate a csr4 file: magic "PCSR", version 4, size, scanner block
 * (the checks and messages of Csr4Matrix::Csr4Matrix, csr4matrix.cpp:36-178),
 * then map the whole data section (Csr4Matrix::mapRows(0, rows),
 * csr4matrix.cpp:200-254).  The size check uses 64-bit arithmetic. */
int mlem_csr4_open(const char *path, mlem_csr4 **out);
int mlem_csr4_close(mlem_csr4 *m);
int mlem_csr4_info(const mlem_csr4 *m, uint32_t *rows, uint32_t *cols,
                   uint64_t *nnz, uint8_t *flags, uint32_t *scan_config_bytes);
int mlem_csr4_scanner(const mlem_csr4 *m, mlem_scanner *out);
/* Inclusive row-end offsets, u64[rows] (csr4matrix.hpp:138-149). */
const uint64_t *mlem_csr4_row_ends(const mlem_csr4 *m);
/* Mapped AoS data: nnz x {uint32 column, float32 value} (matrixelement.hpp:37-46). */
const void *mlem_csr4_data(const mlem_csr4 *m);
/* Csr4Matrix::elementsInRow (csr4matrix.cpp:267-271). */
uint64_t mlem_csr4_elements_in_row(const mlem_csr4 *m, uint32_t row);

/* AoS -> SoA CSR (csr_format_for_cuda, main.cu:518-554) with 64-bit offsets:
 * rowptr u64[rows+1], col u32[nnz], val f32[nnz] (values copied bitwise).
 * threads <= 0 uses the OpenMP default. */
int mlem_csr4_to_csr(const mlem_csr4 *m, uint64_t *rowptr, uint32_t *col,
                     float *val, int threads);

/* Write a csr4 file (the inverse of mlem_csr4_open).  scanner may be NULL
 * (a minimal 60-byte block: 1 layer, blocksize 1, 1 ring, grid cols x 1 x 1). */
int mlem_csr4_write(const char *path, uint32_t rows, uint32_t cols,
                    uint64_t nnz, const uint64_t *rowptr, const uint32_t *col,
                    const float *val, const mlem_scanner *scanner,
                    uint8_t flags);

/* ------------------------------------------------------------------------ */
/* raw vector files (helper_files_common/vector.hpp:33-53, 77-87)           */
/* ------------------------------------------------------------------------ */

/* Read a raw little-endian file of elem_bytes-sized elements; the buffer is
 * malloc'd by the library (free with mlem_host_free).  Fails like
 * Vector(const std::string&) when the file is missing or its size is not a
 * multiple of elem_bytes. */
int mlem_read_raw(const char *path, size_t elem_bytes, void **data, uint64_t *n);
int mlem_write_raw(const char *path, const void *data, size_t elem_bytes,
                   uint64_t n);
void mlem_host_free(void *p);

/* ------------------------------------------------------------------------ */
/* host preprocessing (main.cu:42-143, 518-611; sptrans.h:397-545)          */
/* ------------------------------------------------------------------------ */

/* Column sums in float32 (calcColumnSums, main.cu:558-580), accumulated per
 * column in row order (the one-thread OpenMP order).  norm: f32[cols]. */
int mlem_colsum_f32(const uint64_t *rowptr, const uint32_t *col,
                    const float *val, uint32_t rows, uint32_t cols,
                    float *norm);

/* Initial image (main.cu:83-97): sum_norm = serial f32 sum of norm,
 * sum_g = int32 sum of g[0..rows), init = (float)sum_g / sum_norm,
 * f0[:] = init.  init_out may be NULL. */
int mlem_initial_image(const int32_t *g, uint32_t rows, const float *norm,
                       uint32_t cols, float *f0, float *init_out);

/* CSR(A) -> CSR(A^T) by a stable counting sort, bit-exact with
 * sptrans_scanTrans_specialized (sptrans.h:397-545): colptrT u64[cols+1],
 * rowT u32[nnz], valT f32[nnz] (valT may be NULL, like csrVal == NULL there).
 * The output does not depend on `threads` (<= 0: OpenMP default).  Uses
 * (threads+1) x cols x 8 bytes of temporary memory, no per-nnz temporaries. */
int mlem_transpose(uint32_t rows, uint32_t cols, uint64_t nnz,
                   const uint64_t *rowptr, const uint32_t *col,
                   const float *val, uint64_t *colptrT, uint32_t *rowT,
                   float *valT, int threads);

/* nnz-balanced row partition (partitionMatrix, main.cu:593-611):
 * segments u32[parts+1], segment_rows u32[parts], segment_nnzs u32[parts]
 * (truncated to 32 bits as the reference does; segment_nnzs64 carries the
 * exact count and may be NULL), offsets u64[parts]. */
int mlem_partition(const uint64_t *rowptr, uint64_t nnz, uint32_t rows,
                   uint32_t parts, uint32_t *segments, uint32_t *segment_rows,
                   uint32_t *segment_nnzs, uint64_t *segment_nnzs64,
                   uint64_t *offsets);

/* memcpy of `bytes` on the library's OpenMP threads (<= 0: OpenMP default),
 * which follow OMP_PROC_BIND / OMP_PLACES: the device library's csr4 loader
 * fills its pinned staging buffers with it.  No counterpart in the
 * reference. */
int mlem_host_copy(void *dst, const void *src, uint64_t bytes, int threads);

/* An order of the rows (LORs) of a system matrix for the batched layouts
 * (mlem_matrix_set_row_order in mlem_hip.h): order[i] = place of row i.
 * From the matrix and its image grid (nx, ny, nz with x fastest, as the
 * columns of the csr4 scanner block's grid) alone: each row's weighted voxel
 * moments give its line in the transaxial plane (angle phi, offset s) and
 * its mean slice; rows sort by (10-degree block of phi, s, phi, slice), so
 * that the LORs through one voxel take nearby places.  Deterministic for any
 * `threads` (<= 0: OpenMP default).  No counterpart in the reference. */
int mlem_lor_order(const uint64_t *rowptr, const uint32_t *col,
                   const float *val, uint32_t rows, uint32_t cols,
                   const uint32_t grid[3], int threads, uint32_t *order);

/* ------------------------------------------------------------------------ */
/* synthetic data (the reference's data came from external tools)           */
/* ------------------------------------------------------------------------ */

/* Cylindrical scanner: n_rings axial rings of n_per_ring crystals on a circle
 * of radius `radius` (mm, crystal inner face), crystal size `crystal` =
 * {tangential, axial, depth} (mm); image grid nx x ny x nz voxels of `voxel`
 * (mm), centred on the scanner axis.  Rows are all crystal pairs (a<b) in
 * lexicographic order, crystal id = ring * n_per_ring + k, crystal k at angle
 * 2*pi*k/n_per_ring.  With sub_rays == 0 each LOR is one ray between the
 * crystal centres on the inner face (exact Siddon lengths, used for the 2D
 * ring: nz = 1, n_rings = 1); with sub_rays = S each LOR averages the Siddon
 * lengths of S rays between points drawn uniformly inside both crystal
 * volumes, merged per voxel.  Columns: x + nx*(y + ny*z) (scannerconfig.cpp:31-35). */
typedef struct mlem_geometry {
    uint32_t n_rings;
    uint32_t n_per_ring;
    float radius;
    float crystal[3];
    float ring_pitch;   /* axial distance between ring centres (mm) */
    uint32_t grid[3];
    float voxel[3];
    uint32_t sub_rays;
} mlem_geometry;

typedef struct mlem_csr_host {
    uint32_t rows, cols;
    uint64_t nnz;
    uint64_t *rowptr; /* library-owned; release with mlem_csr_host_free */
    uint32_t *col;
    float *val;
} mlem_csr_host;

uint32_t mlem_geometry_rows(const mlem_geometry *geo);
uint32_t mlem_geometry_cols(const mlem_geometry *geo);

/* Build rows [row_begin, row_end) of the system matrix (row_end = 0: all
 * rows).  Deterministic for a seed, independent of the thread count. */
int mlem_gen_system_matrix(const mlem_geometry *geo, uint64_t seed,
                           uint32_t row_begin, uint32_t row_end, int threads,
                           mlem_csr_host *out);
void mlem_csr_host_free(mlem_csr_host *m);

/* Phantom of n_shapes seeded ellipses (nz == 1) or ellipsoids, activity
 * U(a_min, a_max) each, summed where they overlap; x_true f32[cols]. */
int mlem_gen_phantom(const mlem_geometry *geo, uint64_t seed,
                     uint32_t n_shapes, float a_min, float a_max,
                     float *x_true);

/* y = Poisson(scale * A x) with scale = total_counts / sum(A x), using the
 * library's own counter-based RNG (inversion below lambda 10, Hoermann's PTRS
 * above), one stream per row; y int32[rows], expected_out f64[rows] optional. */
int mlem_gen_sinogram(uint32_t rows, const uint64_t *rowptr,
                      const uint32_t *col, const float *val,
                      const float *x, double total_counts, uint64_t seed,
                      int threads, int32_t *y, double *expected_out);

/* Poisson variate for a given lambda from the generator's stream (seed, idx);
 * exposed for distribution tests. */
int64_t mlem_poisson_sample(double lambda, uint64_t seed, uint64_t idx);

#ifdef __cplusplus
}
#endif
#endif /* MLEM_HOST_H */
