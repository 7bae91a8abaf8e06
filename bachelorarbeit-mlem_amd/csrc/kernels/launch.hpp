// Internal launchers shared by the C ABI (abi.hip, batch.hip, tiles_only.hip) and the kernels
// (spmv.hip, transpose.hip).  All launchers are asynchronous on `s` and
// return hipSuccess or the launch error.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>
#include <vector>

namespace mlem {

constexpr int kWG = 256;              // work-items per workgroup (4 wave64)
constexpr uint32_t kMaxSecsize = 32;  // merge-path items per work-item

// What a finished row sum y_r turns into.
enum Epi : int {
    EPI_PLAIN = 0,   // out[r] = y_r
    EPI_RATIO = 1,   // out[r] = y_r != 0 ? (float)g[r] / y_r : 0      (calcCorrel)
    EPI_UPDATE = 2,  // f[r] = norm[r] == 0 ? f[r]*y_r : f[r]*y_r/norm[r] (calcUpdate)
};

struct EpiArgs {
    float *out;                // PLAIN / RATIO output
    const int32_t *g;          // RATIO counts
    const float *norm;         // UPDATE sensitivity
    float *f;                  // UPDATE image, in place
    const uint32_t *active;    // blocked UPDATE over several right-hand sides: [BC] flags, NULL = all
};

// Tile partition of the merge path for one matrix and one secsize.
struct MergePlan {
    uint32_t rows = 0;
    uint64_t nnz = 0;
    uint32_t secsize = 0;
    uint64_t tile_items = 0;  // kWG * secsize
    uint32_t ntiles = 0;
};

MergePlan merge_plan(uint32_t rows, uint64_t nnz, uint32_t secsize);
// Workspace split: tile_row u32[ntiles+1], tile_nz u64[ntiles+1] (the
// partition, matrix-dependent only) and carry/head f32[ntiles] (per call).
size_t merge_partition_bytes(const MergePlan &p);
size_t merge_scratch_bytes(const MergePlan &p);

hipError_t launch_merge_partition(const MergePlan &p, const uint64_t *rowptr, void *partition, hipStream_t s);
hipError_t launch_merge_spmv(const MergePlan &p, const void *partition, void *scratch, const uint64_t *rowptr,
                             const uint32_t *col, const float *val, const float *x, Epi epi, const EpiArgs &ea,
                             hipStream_t s);

uint32_t auto_lanes(uint32_t rows, uint64_t nnz);
bool valid_lanes(uint32_t lanes);
hipError_t launch_csr_vector(uint32_t lanes, const uint64_t *rowptr, const uint32_t *col, const float *val,
                             const float *x, uint32_t rows, Epi epi, const EpiArgs &ea, hipStream_t s);
hipError_t launch_scatter(uint32_t lanes, const uint64_t *rowptr, const uint32_t *col, const float *val,
                          const float *xrow, float *out, uint32_t rows, hipStream_t s);
hipError_t launch_ratio(const int32_t *g, float *fwproj, uint32_t rows, hipStream_t s);
// In place over bwproj: bwproj = update(f, bwproj, norm)  (calcUpdate)
hipError_t launch_update(const float *f, const float *norm, float *bwproj, uint32_t cols, hipStream_t s);
// f = update(f, bp, norm) and bp = 0 (the scatter path's update + re-zero)
hipError_t launch_update_image(float *f, const float *norm, float *bp, uint32_t cols, hipStream_t s);
hipError_t launch_fill(float *dst, float v, uint64_t n, hipStream_t s);
hipError_t launch_fill_f64(double *dst, double v, uint64_t n, hipStream_t s);
hipError_t launch_fill_u32(uint32_t *dst, uint32_t v, uint64_t n, hipStream_t s);
// csr4 AoS entries [first, first+n) -> col/val; bit 1 of *bad when a
// column is >= cols.
hipError_t launch_aos_split(const void *aos, uint64_t n, uint64_t first, uint32_t cols, uint32_t *col, float *val,
                            uint32_t *bad, hipStream_t s);
// *first = min(*first, smallest k < n with col[k] >= cols).
hipError_t launch_first_bad_column(const uint32_t *col, uint64_t n, uint32_t cols, uint64_t *first, hipStream_t s);
// 4-byte elements: dst[i*stride] = src[i] / dst[i] = src[i*stride], i < n.
hipError_t launch_strided_put(void *dst, uint32_t stride, const void *src, uint64_t n, hipStream_t s);
// 4-byte elements in row order: dst[order[i]*stride] = src[i], i < n.
hipError_t launch_ordered_put(void *dst, uint32_t stride, const void *src, const uint32_t *order, uint32_t n,
                              hipStream_t s);
// The matrix with row i of A in place order[i] (a permutation of the rows):
// rowptr_out u64[rows+1], col_out / val_out [nnz], each row's entries as in
// A.  Allocates its scratch and waits for the stream (reorder.hip).
hipError_t reorder_rows(uint32_t rows, uint64_t nnz, const uint64_t *rowptr, const uint32_t *col, const float *val,
                        const uint32_t *order, uint64_t *rowptr_out, uint32_t *col_out, float *val_out,
                        hipStream_t s);
hipError_t launch_strided_get(void *dst, const void *src, uint32_t stride, uint64_t n, hipStream_t s);
// dst = src (16-byte aligned), non-temporal streaming copy.
hipError_t launch_copy(float *dst, const float *src, uint64_t n, hipStream_t s);
hipError_t launch_colsum_rows_f64(const uint64_t *rowptrT, const float *valT, uint32_t cols, double *s64, float *s32,
                                  hipStream_t s);
hipError_t launch_colsum_scatter_f64(const uint64_t *rowptr, const uint32_t *col, const float *val, uint32_t rows,
                                     double *s64, hipStream_t s);
// s[c] = sum_k valT[k] * w[rowT[k]] over row c of A^T, in FP64.
hipError_t launch_colsum_weighted_f64(const uint64_t *rowptrT, const uint32_t *rowT, const float *valT, const float *w,
                                      uint32_t cols, double *s64, float *s32, hipStream_t s);
hipError_t launch_f64_to_f32(const double *a, float *b, uint32_t n, hipStream_t s);
// Poisson log-likelihood partial sums over rows of A x: per-block FP64
// partials into part[grid], grid = loglik_blocks(rows).
uint32_t loglik_blocks(uint32_t rows);
hipError_t launch_loglik(const uint64_t *rowptr, const uint32_t *col, const float *val, const float *x,
                         const int32_t *g, uint32_t rows, double *part, hipStream_t s);
// The same from y = A x computed elsewhere (the tiled forward projection).
hipError_t launch_loglik_y(const float *y, const int32_t *g, uint32_t rows, double *part, hipStream_t s);

// Blocked SpMV (tiled.hip): tiles of R rows x W columns with the W-wide
// segment of x staged in LDS, entries packed to 6 bytes in units of
// kTileUnit.  The units of each row block are split into runs of nearly
// equal length, one run per workgroup; a workgroup writes the partial sums
// of its block's rows, a reduction kernel adds the runs of each block in
// order and applies the epilogue.
constexpr uint32_t kTileXPad = 8192;     // pad gathered vectors to a multiple of this
constexpr uint32_t kTileSpanMax = 256;   // most tiles one run may touch
constexpr uint32_t kTileE = 8;           // entries per work-item of a unit
constexpr uint32_t kTileUnit = kWG * kTileE;
// Stored value type of a tiled matrix: f32, or bf16 / fp16 holding
// value / scale[row] with an f32 scale per row (the largest |value|).
// Stored values of the blocked kernels: f32, or 16 bits with a per-row f32
// scale: BF16, FP16, or UF16, an unsigned 16-bit float for nonnegative
// matrices (no sign bit: 4 exponent bits, 12 mantissa bits; value / row
// maximum in [2^-15, 1], no zero: code 0 is 2^-15, padding multiplies a zero
// x entry).
enum ValType : uint32_t { VAL_F32 = 0, VAL_BF16 = 1, VAL_F16 = 2, VAL_UF16 = 3 };
constexpr bool is_uf16(uint32_t vt) { return vt == VAL_UF16; }
constexpr int kUf16ExpBits = 4;
// The per-row scale of a 16-bit type: this times the row's largest |value|.
inline float value_scale_mult(uint32_t vt) { return vt == VAL_F16 ? 1.0f / 16384.0f : 1.0f; }
struct TilePlan {
    uint32_t rows = 0, cols = 0;
    uint64_t nnz = 0;
    uint32_t W = 0, R = 0, nseg = 0, nblk = 0;
    uint64_t ntiles = 0;
    uint64_t cap = 0;  // entries including padding (set by launch_tile_count)
    uint32_t vt = VAL_F32;
    // Group layout: false = row steps (a group of 8 may hold several rows,
    // the single right-hand-side kernels), true = quads (every 4 entries
    // belong to one row, the several right-hand-side kernels).
    bool quad = false;
    // 16-bit values that are codes already (the transposed tiles hold the
    // codes of A's tiles, so both sides use one quantized matrix): scale 1.
    bool unit_scale = false;
    uint32_t vb() const { return vt == VAL_F32 ? 4u : 2u; }  // bytes per stored value
};
struct TiledMatrix {
    uint64_t *tile_off = nullptr;  // [ntiles+1] entry offsets (padded to 8)
    uint32_t *tile_nz = nullptr;   // [ntiles] entries including padding
    uint8_t *tdata = nullptr;      // vb+2 bytes per entry: per unit, values then group words
    float *scale = nullptr;        // [rows] value scales (16-bit values only)
};
// Work list on the host: 4 u32 per run {row block, first tile (index into
// tiles), first unit in that tile, units}; 4 u32 per non-empty tile
// {segment, entries, entry offset lo, hi}; wl_first[b] = first run of
// row block b (nblk+1 entries).
struct TileWorkList {
    uint32_t span = 1, nwg = 0, ntl = 0;
    uint32_t bad = 0;  // quad tiles not aligned as the stream kernel needs
    std::vector<uint32_t> wl, tiles, wl_first;
    std::vector<uint32_t> order;  // run of each workgroup, in launch order (column-segment order)
};
// The same on the device.
struct TileWork {
    uint32_t span = 1, nwg = 0, ntl = 0;
    uint32_t max_runs = 0;  // the most runs of one row block
    const uint4 *wl = nullptr;
    const uint4 *tiles = nullptr;
    const uint32_t *wl_first = nullptr;
    const uint32_t *order = nullptr;  // run of each workgroup (null: workgroup w takes run w)
};
TilePlan tile_plan(uint32_t rows, uint32_t cols, uint64_t nnz, uint32_t W, uint32_t R, uint32_t vt, bool quad);
// The 16-bit quantization of a CSR's values as the tiles store it: per-row
// scale (largest |value|, times 2^-14 for FP16), codes rounded to nearest
// (q, 2 bytes per nonzero of scratch) and decoded to f32 without the scale.
hipError_t launch_quantize_codes(const uint64_t *rowptr, const float *val, uint32_t rows, uint64_t nnz, uint32_t vt,
                                 float *scale, float *codes, uint16_t *q, hipStream_t s);
// The same scales, and each value replaced by its decoded code in place (no
// code scratch); *bad (device, may be null) bit 8: a negative value for uf16.
hipError_t launch_quantize_in_place(const uint64_t *rowptr, float *val, uint32_t rows, uint32_t vt, float *scale,
                                    uint32_t *bad, hipStream_t s);
bool valid_tile_width(uint32_t W);
// Kernel shapes built: (W, right-hand sides per pass) in {1024, 2048} x 1,
// {512, 1024} x 4, 512 x 8.
bool valid_tile_config(uint32_t W, uint32_t BC);
bool valid_tile_rows(uint32_t R);
size_t tiled_matrix_bytes(const TilePlan &p);
size_t tiled_part_bytes(const TilePlan &p, uint32_t nwg, uint32_t BC);
size_t tiled_lds_bytes(const TilePlan &p, uint32_t span, uint32_t BC);
size_t tile_count_scratch_bytes(const TilePlan &p);
size_t tile_fill_scratch_bytes(const TilePlan &p);  // after the count (uses p.cap)
TiledMatrix tiled_matrix_view(const TilePlan &p, void *mem);
// Building the tiles from a CSR on the device takes two calls, each waits.
// Count: sizes and offsets into `scratch`, sets p.cap; *flags_host bit 1 =
// a row is not sorted by column (then p.cap is not set and the tiles must
// not be filled).  Fill: the tiled arrays into `mem` (tiled_matrix_bytes(p))
// with the same scratch; *flags_host bit 4 = internal layout error.
hipError_t launch_tile_count(TilePlan &p, const uint64_t *rowptr, const uint32_t *col, void *scratch,
                             uint32_t *flags_host, hipStream_t s);
hipError_t launch_tile_fill(const TilePlan &p, const uint64_t *rowptr, const uint32_t *col, const float *val,
                            void *mem, void *scratch, void *fill_scratch, uint32_t *flags_host, hipStream_t s);
// The same into arrays given by a view: tile_off receives the count's
// offsets (relative to tdata), tile_nz the sizes.  Tiles built in row
// chunks (whole row blocks, each chunk a CSR of its own) fill the chunk's
// part of the whole matrix's tiles: tdata shifted by the chunk's first entry
// offset, tile_nz by its first tile, tile_off a scratch of ntiles + 1.
hipError_t launch_tile_fill_into(const TilePlan &p, const uint64_t *rowptr, const uint32_t *col, const float *val,
                                 const TiledMatrix &tm, void *scratch, void *fill_scratch, uint32_t *flags_host,
                                 hipStream_t s);
// After launch_tile_count: the tile offsets (ntiles + 1) and sizes (ntiles)
// it left in `scratch` (device pointers).
const uint64_t *tile_count_offsets(const TilePlan &p, void *scratch);
const uint32_t *tile_count_sizes(const TilePlan &p, void *scratch);
// Runs for about target_wgs workgroups from the host copies of tile_nz and
// tile_off.
TileWorkList tile_work_list(const TilePlan &p, const uint32_t *tile_nz, const uint64_t *tile_off,
                            uint32_t target_wgs);
// Y = A X for BC right-hand sides in two launches: the partial row sums of
// every run into part (tiled_part_bytes(p, wk.nwg, BC)), then their
// reduction with the row scales and the epilogue.  Vectors are [n][BC]
// (right-hand side fastest); X must be readable (and finite) up to
// nseg * W rows: callers pad their vectors to whole segments (kTileXPad).
// nt: work-items per workgroup of the kernel for several right-hand sides
// (128, 256 or 512; 0 = 512), each taking 512 / nt groups of a window.
hipError_t launch_tiled_partials(const TilePlan &p, const TiledMatrix &tm, const TileWork &wk, const float *x,
                                 float *part, uint32_t BC, uint32_t nt, hipStream_t s);
bool valid_stream_threads(uint32_t nt);
hipError_t launch_tiled_reduce(const TilePlan &p, const TiledMatrix &tm, const TileWork &wk, const float *part,
                               uint32_t BC, Epi epi, const EpiArgs &ea, hipStream_t s);

// Sets the message of mlem_last_error() (this thread).
void set_last_error(const char *msg);

// Device transpose (stable counting sort): see transpose.hip.
size_t transpose_scratch_bytes(uint32_t rows, uint32_t cols, uint64_t nnz);
hipError_t launch_transpose(uint32_t rows, uint32_t cols, uint64_t nnz, const uint64_t *rowptr, const uint32_t *col,
                            const float *val, uint64_t *colptrT, uint32_t *rowT, float *valT, void *scratch,
                            hipStream_t s);
// The same result with at most `budget` bytes of scratch (allocated and freed
// here; the call waits): column ranges sorted one at a time.  *ranges = how
// many; hipErrorOutOfMemory when one column alone does not fit.
hipError_t transpose_in_ranges(uint32_t rows, uint32_t cols, uint64_t nnz, const uint64_t *rowptr,
                               const uint32_t *col, const float *val, uint64_t *colptrT, uint32_t *rowT, float *valT,
                               size_t budget, uint32_t *ranges, hipStream_t s);
// The same with the ranges' columns a multiple of `align` (but the last)
// and, when `sink` is set, each range handed to it as a slab of A^T's rows
// [c0, c1) (n entries from offset off0 = colptrT[c0] of A^T: rowT and valT
// of the slab, its row pointers colptrT[c0..c1] minus off0) instead of
// being written to rowT / valT (then null).  colptrT must still hold cols + 1
// entries.  The slab's arrays are valid until the sink returns.
using RangeSink = std::function<hipError_t(uint32_t c0, uint32_t c1, uint64_t off0, uint64_t n,
                                           const uint32_t *rowT, const float *valT)>;
hipError_t transpose_ranges_to(uint32_t rows, uint32_t cols, uint64_t nnz, const uint64_t *rowptr,
                               const uint32_t *col, const float *val, uint64_t *colptrT, uint32_t *rowT, float *valT,
                               size_t budget, uint32_t align, const RangeSink &sink, uint32_t *ranges_out,
                               hipStream_t s);
// dst[i] = src[i] - base for i = 0..n.
hipError_t launch_rebase(const uint64_t *src, uint64_t base, uint32_t n, uint64_t *dst, hipStream_t s);

}  // namespace mlem
