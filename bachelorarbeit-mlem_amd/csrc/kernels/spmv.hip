// SpMV kernels of the MLEM iteration for CDNA4 (gfx950, wave64).
//
//  * csr_vector_kernel<G, EPI>: a group of G work-items per row
//    (calcFwProj_csr_vector / calcBwProj_csr_vector, kernel.cu:118-153,
//    297-336).  Consecutive 4-byte non-temporal loads of col/val, eight
//    rounds in flight, fixed-order xor shuffle reduction, plain store
//    through the epilogue.
//  * merge_*: block-level merge-path SpMV (calcFwProj_merge_based /
//    calcBwProj_merge_based, kernel.cu:52-105,190-286).  The merge path of
//    (row ends, nonzeros) is cut into tiles of 256*S items; each workgroup
//    stages its tile's products and row ends in LDS, every work-item walks S
//    items, a segmented scan over the workgroup joins rows split between
//    work-items, and rows split between tiles are finished by a fix-up kernel
//    that adds the tile carries in tile order.  No float atomics: results are
//    bitwise reproducible.
//  * scatter_kernel<G>: BP without a transpose (calcBwProj_none_trans,
//    kernel.cu:165-177,347-370), float atomics into L2.
//  * elementwise ratio / update / fill, FP64 column sums, FP64 log-likelihood.
#include <algorithm>

#include "launch.hpp"

namespace mlem {
namespace {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Streamed matrix data is read once per pass: non-temporal loads keep it from
// evicting the gathered image / ratio vector from L2.
template <typename T>
__device__ __forceinline__ T ldnt(const T *p) {
    return __builtin_nontemporal_load(p);
}

// The epilogue of row r with row sum y: calcCorrel (kernel.cu:16-18), g / y
// where y != 0, else y stays; calcUpdate (kernel.cu:33-37), (f * y) / norm,
// or f * y where norm == 0; or the plain sum.  epi_load reads the row's
// inputs (RATIO: the count; UPDATE: the image value and the sensitivity) and
// emit_with forms the result, so that a kernel can issue those reads before
// its row sum and their memory round trip overlaps the sum's.
struct EpiIn {
    float a, b;
};

template <int EPI>
__device__ __forceinline__ EpiIn epi_load(uint32_t r, const EpiArgs &ea) {
    if constexpr (EPI == EPI_RATIO) return {(float)ea.g[r], 0.0f};
    else if constexpr (EPI == EPI_UPDATE) return {ea.f[r], ea.norm[r]};
    else return {0.0f, 0.0f};
}

template <int EPI>
__device__ __forceinline__ void emit_with(uint32_t r, float y, const EpiIn &in, const EpiArgs &ea) {
    if constexpr (EPI == EPI_PLAIN) {
        ea.out[r] = y;
    } else if constexpr (EPI == EPI_RATIO) {
        ea.out[r] = (y != 0.0f) ? in.a / y : y;
    } else {
        const float v = in.a * y;
        ea.f[r] = (in.b == 0.0f) ? v : v / in.b;
    }
}

template <int EPI>
__device__ __forceinline__ void emit(uint32_t r, float y, const EpiArgs &ea) {
    emit_with<EPI>(r, y, epi_load<EPI>(r, ea), ea);
}

// ---------------------------------------------------------------------------
// CSR-vector
// ---------------------------------------------------------------------------
template <int G>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

template <int G>
__device__ __forceinline__ double group_sum_f64(double v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Row sum of val[k] * x[col[k]] over [b, e) split over the G work-items of a
// group (this work-item is `lane`); per-lane partials, not yet reduced.
// Consecutive work-items take consecutive entries (4-byte loads), so one
// gather instruction reads neighbouring columns, which share cache lines.
// Rounds of U = 8 entries per work-item are loaded before any is used; the
// last (partial) round is issued the same way with the entries past the row
// end masked off, so a short row costs one load latency, not one per entry.
template <int G>
__device__ __forceinline__ float row_partial(const uint32_t *__restrict__ col, const float *__restrict__ val,
                                             const float *__restrict__ x, uint64_t b, uint64_t e, uint32_t lane) {
    constexpr int U = 8;
    float acc = 0.0f;
    uint64_t k = b + lane;
    for (; k + (uint64_t)(U - 1) * G < e; k += (uint64_t)U * G) {
        uint32_t c[U];
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            c[u] = ldnt(col + k + (uint64_t)u * G);
            v[u] = ldnt(val + k + (uint64_t)u * G);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) acc += v[u] * x[c[u]];
    }
    if (k < e) {  // fewer than U rounds left
        uint32_t c[U];
        float v[U];
        bool in[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            in[u] = k + (uint64_t)u * G < e;
            c[u] = in[u] ? ldnt(col + k + (uint64_t)u * G) : 0u;
            v[u] = in[u] ? ldnt(val + k + (uint64_t)u * G) : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (in[u]) acc += v[u] * x[c[u]];
    }
    return acc;
}

template <int G, int EPI>
__global__ __launch_bounds__(kWG) void csr_vector_kernel(const uint64_t *__restrict__ rowptr,
                                                         const uint32_t *__restrict__ col,
                                                         const float *__restrict__ val, const float *__restrict__ x,
                                                         uint32_t rows, EpiArgs ea) {
    const uint32_t row = blockIdx.x * (uint32_t)(kWG / G) + threadIdx.x / G;
    const uint32_t lane = threadIdx.x % G;
    if (row >= rows) return;  // whole groups leave together
    const EpiIn in = epi_load<EPI>(row, ea);  // every work-item of the group: one address, one request
    const float acc = group_sum<G>(row_partial<G>(col, val, x, rowptr[row], rowptr[row + 1], lane));
    if (lane == 0) emit_with<EPI>(row, acc, in, ea);
}

// ---------------------------------------------------------------------------
// merge path
// ---------------------------------------------------------------------------

// Merge of the row ends a_i = rowptr[i+1] (i < rows) with the nonzero
// positions 0..nnz-1, a_i taken before nonzero j iff a_i <= j (the order of
// merge_based_work, kernel.cu:271-282).  The split point of diagonal d is the
// smallest i in [max(0, d-nnz), min(d, rows)] with rowptr[i+1] + i >= d,
// i.e. rowptr[i+1] > (d-i) - 1 (the search of merge_based_start,
// kernel.cu:202-229, without its read of rowptr[rows+1]).
__global__ void merge_partition_kernel(const uint64_t *__restrict__ rowptr, uint32_t rows, uint64_t nnz,
                                       uint64_t tile_items, uint32_t ntiles, uint32_t *__restrict__ tile_row,
                                       uint64_t *__restrict__ tile_nz) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntiles) return;
    const uint64_t L = (uint64_t)rows + nnz;
    uint64_t d = (uint64_t)t * tile_items;
    if (d > L) d = L;
    uint64_t lo = d > nnz ? d - nnz : 0;
    uint64_t hi = d < rows ? d : rows;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (rowptr[mid + 1] + mid >= d)
            hi = mid;
        else
            lo = mid + 1;
    }
    tile_row[t] = (uint32_t)lo;
    tile_nz[t] = d - lo;
}

template <int EPI>
__global__ __launch_bounds__(kWG) void merge_spmv_kernel(const uint64_t *__restrict__ rowptr,
                                                         const uint32_t *__restrict__ col,
                                                         const float *__restrict__ val, const float *__restrict__ x,
                                                         uint32_t S, const uint32_t *__restrict__ tile_row,
                                                         const uint64_t *__restrict__ tile_nz,
                                                         float *__restrict__ tile_carry,
                                                         float *__restrict__ tile_head, EpiArgs ea) {
    extern __shared__ __attribute__((aligned(16))) float smem[];
    __shared__ float s_wv[kWG / 64];
    __shared__ int s_wf[kWG / 64];

    const uint32_t t = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    const uint32_t i0 = tile_row[t], i1 = tile_row[t + 1];
    const uint64_t j0 = tile_nz[t], j1 = tile_nz[t + 1];
    const uint32_t nr = i1 - i0;              // row ends in this tile
    const uint32_t nz = (uint32_t)(j1 - j0);  // nonzeros in this tile
    const uint32_t h = (uint32_t)(j0 & 3u);   // s_val[h + jl] holds nonzero j0 + jl
    float *s_val = smem;
    uint32_t *s_end = reinterpret_cast<uint32_t *>(smem + ((h + nz + 3u) & ~3u));

    // Stage products val*x[col] with aligned 16-byte loads (partial quads at
    // the tile edges element by element) and the local row ends.
    const uint64_t ja = j0 - h;
    const uint32_t nq = (h + nz + 3u) >> 2;
    for (uint32_t q = tid; q < nq; q += kWG) {
        const uint64_t jq = ja + 4ull * q;
        if (jq >= j0 && jq + 4 <= j1) {
            const u32x4 c = ldnt(reinterpret_cast<const u32x4 *>(col + jq));
            const f32x4 v = ldnt(reinterpret_cast<const f32x4 *>(val + jq));
            f32x4 p;
            p.x = v.x * x[c.x];
            p.y = v.y * x[c.y];
            p.z = v.z * x[c.z];
            p.w = v.w * x[c.w];
            *reinterpret_cast<f32x4 *>(s_val + 4u * q) = p;
        } else {
            for (uint32_t e = 0; e < 4; ++e) {
                const uint64_t j = jq + e;
                if (j >= j0 && j < j1) s_val[4u * q + e] = val[j] * x[col[j]];
            }
        }
    }
    for (uint32_t q = tid; q < nr; q += kWG) s_end[q] = (uint32_t)(rowptr[(uint64_t)i0 + 1 + q] - j0);
    __syncthreads();

    // Each work-item walks S consecutive merge items of the tile.
    const uint32_t Tt = nr + nz;
    uint32_t d0 = tid * S;
    if (d0 > Tt) d0 = Tt;
    const uint32_t d1 = (d0 + S < Tt) ? d0 + S : Tt;
    uint32_t lo = d0 > nz ? d0 - nz : 0u;
    uint32_t hi = d0 < nr ? d0 : nr;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_end[mid] + mid >= d0)
            hi = mid;
        else
            lo = mid + 1;
    }
    const uint32_t first = lo;
    uint32_t i = lo, j = d0 - lo;
    float acc = 0.0f, head = 0.0f;
    int has_end = 0;
    for (uint32_t d = d0; d < d1; ++d) {
        if (i < nr && s_end[i] <= j) {
            if (has_end)
                emit<EPI>(i0 + i, acc, ea);  // row started and ended in this work-item
            else {
                head = acc;
                has_end = 1;
            }
            acc = 0.0f;
            ++i;
        } else {
            acc += s_val[h + j];
            ++j;
        }
    }

    // Segmented inclusive scan of (has_end, carry) over the workgroup:
    // (f1,v1) (+) (f2,v2) = (f1|f2, f2 ? v2 : v1+v2), left operand first.
    const uint32_t lane = tid & 63u, w = tid >> 6;
    float v = acc;
    int f = has_end;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const float vu = __shfl_up(v, o);
        const int fu = __shfl_up(f, o);
        if (lane >= (uint32_t)o) {
            if (!f) v = vu + v;
            f |= fu;
        }
    }
    if (lane == 63) {
        s_wv[w] = v;
        s_wf[w] = f;
    }
    __syncthreads();
    float pv = 0.0f;
    int pf = 0;
    for (uint32_t u = 0; u < w; ++u) {
        pv = s_wf[u] ? s_wv[u] : pv + s_wv[u];
        pf |= s_wf[u];
    }
    const float iv = f ? v : pv + v;
    const int ifl = pf | f;
    float ev = __shfl_up(iv, 1);
    int ef = __shfl_up(ifl, 1);
    if (lane == 0) {
        ev = pv;
        ef = pf;
    }
    if (has_end) {
        const float tot = ev + head;
        // The tile's first finished row started in an earlier tile: leave its
        // partial to the fix-up kernel.
        if (!ef && rowptr[i0] < j0)
            tile_head[t] = tot;
        else
            emit<EPI>(i0 + first, tot, ea);
    }
    if (tid == kWG - 1) tile_carry[t] = iv;
}

// Rows whose nonzeros span tiles: sum the carries of the tiles before the
// one that ends the row, in tile order, then that tile's head partial.
template <int EPI>
__global__ void merge_fixup_kernel(const uint64_t *__restrict__ rowptr, uint32_t ntiles,
                                   const uint32_t *__restrict__ tile_row, const uint64_t *__restrict__ tile_nz,
                                   const float *__restrict__ tile_carry, const float *__restrict__ tile_head,
                                   EpiArgs ea) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= ntiles) return;
    const uint32_t r = tile_row[u];
    if (tile_row[u + 1] == r) return;       // no row ends in tile u
    if (!(rowptr[r] < tile_nz[u])) return;  // its first row starts inside tile u (always so for u == 0)
    uint32_t s = u - 1;
    while (s > 0 && tile_row[s] == r) --s;
    float tot = 0.0f;
    for (uint32_t q = s; q < u; ++q) tot += tile_carry[q];
    emit<EPI>(r, tot + tile_head[u], ea);
}

// ---------------------------------------------------------------------------
// scatter BP, elementwise, sensitivity, log-likelihood
// ---------------------------------------------------------------------------
template <int G>
__global__ __launch_bounds__(kWG) void scatter_kernel(const uint64_t *__restrict__ rowptr,
                                                      const uint32_t *__restrict__ col,
                                                      const float *__restrict__ val, const float *__restrict__ xrow,
                                                      float *__restrict__ out, uint32_t rows) {
    const uint32_t row = blockIdx.x * (uint32_t)(kWG / G) + threadIdx.x / G;
    const uint32_t lane = threadIdx.x % G;
    if (row >= rows) return;
    const float r = xrow[row];
    if (r == 0.0f) return;  // adding +0 leaves every (non-negative zero) sum unchanged
    const uint64_t b = rowptr[row], e = rowptr[row + 1];
    for (uint64_t k = b + lane; k < e; k += G) unsafeAtomicAdd(out + ldnt(col + k), ldnt(val + k) * r);
}

__global__ void ratio_kernel(const int32_t *__restrict__ g, float *__restrict__ fw, uint32_t rows) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < rows) {
        const float y = fw[i];
        if (y != 0.0f) fw[i] = (float)g[i] / y;
    }
}

__global__ void update_kernel(const float *__restrict__ f, const float *__restrict__ norm, float *__restrict__ bw,
                              uint32_t cols) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cols) {
        const float v = f[i] * bw[i];
        const float n = norm[i];
        bw[i] = (n == 0.0f) ? v : v / n;
    }
}

__global__ void update_image_kernel(float *__restrict__ f, const float *__restrict__ norm, float *__restrict__ bp,
                                    uint32_t cols) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cols) {
        const float v = f[i] * bp[i];
        const float n = norm[i];
        f[i] = (n == 0.0f) ? v : v / n;
        bp[i] = 0.0f;
    }
}

template <typename T>
__global__ void fill_kernel(T *__restrict__ dst, T v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = v;
}

// dst = src over n floats, 16 bytes per work-item per step.
__global__ __launch_bounds__(kWG) void copy_kernel(float *__restrict__ dst, const float *__restrict__ src,
                                                   uint64_t n) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const uint64_t n4 = n / 4, stride = (uint64_t)gridDim.x * kWG;
    for (uint64_t i = blockIdx.x * (uint64_t)kWG + threadIdx.x; i < n4; i += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(reinterpret_cast<const f4 *>(src) + i),
                                    reinterpret_cast<f4 *>(dst) + i);
    for (uint64_t i = 4 * n4 + blockIdx.x * (uint64_t)kWG + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// csr4 AoS {u32 col, f32 val} entries [first, first+n) of the matrix -> SoA;
// bit 1 of *bad: a column index >= cols.
__global__ __launch_bounds__(kWG) void aos_split_kernel(const uint32_t *__restrict__ aos, uint64_t n, uint64_t first,
                                                        uint32_t cols, uint32_t *__restrict__ col,
                                                        float *__restrict__ val, uint32_t *__restrict__ bad) {
    const uint64_t stride = (uint64_t)gridDim.x * kWG;
    uint32_t flag = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)kWG + threadIdx.x; i < n; i += stride) {
        const uint32_t c = __builtin_nontemporal_load(aos + 2 * i);
        const uint32_t v = __builtin_nontemporal_load(aos + 2 * i + 1);
        flag |= c >= cols ? 1u : 0u;
        col[first + i] = c;
        val[first + i] = __uint_as_float(v);
    }
    if (flag) atomicOr(bad, 1u);
}

// *first = the smallest k with col[k] >= cols (unchanged when there is none;
// the caller starts it at UINT64_MAX).
__global__ void first_bad_column_kernel(const uint32_t *__restrict__ col, uint64_t n, uint32_t cols,
                                        unsigned long long *__restrict__ first) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x)
        if (col[k] >= cols) {
            atomicMin(first, (unsigned long long)k);
            return;
        }
}

// One slot of an interleaved [n][stride] array <-> a contiguous vector.
template <typename T>
__global__ void strided_put_kernel(T *__restrict__ dst, uint32_t stride, const T *__restrict__ src, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i * stride] = src[i];
}
template <typename T>
__global__ void strided_get_kernel(T *__restrict__ dst, const T *__restrict__ src, uint32_t stride, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i * stride];
}

__global__ void f64_to_f32_kernel(const double *__restrict__ a, float *__restrict__ b, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

// Row sums of A^T (= column sums of A) in FP64, 64 work-items per row.
__global__ __launch_bounds__(kWG) void colsum_rows_kernel(const uint64_t *__restrict__ rowptrT,
                                                          const float *__restrict__ valT, uint32_t cols,
                                                          double *__restrict__ s64, float *__restrict__ s32) {
    const uint32_t row = blockIdx.x * (uint32_t)(kWG / 64) + threadIdx.x / 64;
    const uint32_t lane = threadIdx.x % 64;
    if (row >= cols) return;
    double acc = 0.0;
    for (uint64_t k = rowptrT[row] + lane; k < rowptrT[row + 1]; k += 64) acc += (double)valT[k];
    acc = group_sum_f64<64>(acc);
    if (lane == 0) {
        if (s64) s64[row] = acc;
        if (s32) s32[row] = (float)acc;
    }
}

// s[c] = sum over row c of A^T of valT[k] * w[rowT[k]] (FP64): the column
// sums of a matrix given as row weights times codes.
__global__ __launch_bounds__(kWG) void colsum_weighted_kernel(const uint64_t *__restrict__ rowptrT,
                                                              const uint32_t *__restrict__ rowT,
                                                              const float *__restrict__ valT,
                                                              const float *__restrict__ w, uint32_t cols,
                                                              double *__restrict__ s64, float *__restrict__ s32) {
    const uint32_t row = blockIdx.x * (uint32_t)(kWG / 64) + threadIdx.x / 64;
    const uint32_t lane = threadIdx.x % 64;
    if (row >= cols) return;
    double acc = 0.0;
    for (uint64_t k = rowptrT[row] + lane; k < rowptrT[row + 1]; k += 64) acc += (double)valT[k] * (double)w[rowT[k]];
    acc = group_sum_f64<64>(acc);
    if (lane == 0) {
        if (s64) s64[row] = acc;
        if (s32) s32[row] = (float)acc;
    }
}

__global__ __launch_bounds__(kWG) void colsum_scatter_kernel(const uint64_t *__restrict__ rowptr,
                                                             const uint32_t *__restrict__ col,
                                                             const float *__restrict__ val, uint32_t rows,
                                                             double *__restrict__ s64) {
    const uint32_t row = blockIdx.x * (uint32_t)(kWG / 64) + threadIdx.x / 64;
    const uint32_t lane = threadIdx.x % 64;
    if (row >= rows) return;
    for (uint64_t k = rowptr[row] + lane; k < rowptr[row + 1]; k += 64) unsafeAtomicAdd(s64 + col[k], (double)val[k]);
}

// Per-row l_i = g_i log(y_i) - y_i with y = A x (terms with y_i <= 0
// contribute -y_i), FP64, one partial per workgroup in a fixed order.
template <int G>
__global__ __launch_bounds__(kWG) void loglik_kernel(const uint64_t *__restrict__ rowptr,
                                                     const uint32_t *__restrict__ col, const float *__restrict__ val,
                                                     const float *__restrict__ x, const int32_t *__restrict__ g,
                                                     uint32_t rows, double *__restrict__ part) {
    __shared__ double s[kWG / G];
    const uint32_t row = blockIdx.x * (uint32_t)(kWG / G) + threadIdx.x / G;
    const uint32_t lane = threadIdx.x % G;
    double l = 0.0;
    if (row < rows) {
        const float y = group_sum<G>(row_partial<G>(col, val, x, rowptr[row], rowptr[row + 1], lane));
        const double yd = (double)y;
        l = (yd > 0.0) ? (double)g[row] * log(yd) - yd : -yd;
    }
    if (lane == 0) s[threadIdx.x / G] = l;
    __syncthreads();
    if (threadIdx.x == 0) {
        double acc = 0.0;
        for (int q = 0; q < kWG / G; ++q) acc += s[q];
        part[blockIdx.x] = acc;
    }
}

inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

template <int EPI>
hipError_t launch_merge_epi(const MergePlan &p, const uint32_t *tile_row, const uint64_t *tile_nz, float *carry,
                            float *head, const uint64_t *rowptr, const uint32_t *col, const float *val,
                            const float *x, const EpiArgs &ea, hipStream_t s) {
    const size_t lds = (p.tile_items + 8) * sizeof(float);
    hipLaunchKernelGGL(merge_spmv_kernel<EPI>, dim3(p.ntiles), dim3(kWG), lds, s, rowptr, col, val, x, p.secsize,
                       tile_row, tile_nz, carry, head, ea);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(merge_fixup_kernel<EPI>, dim3(blocks_for(p.ntiles, kWG)), dim3(kWG), 0, s, rowptr, p.ntiles,
                       tile_row, tile_nz, carry, head, ea);
    return hipGetLastError();
}

template <int G>
hipError_t launch_csr_vector_g(const uint64_t *rowptr, const uint32_t *col, const float *val, const float *x,
                               uint32_t rows, Epi epi, const EpiArgs &ea, hipStream_t s) {
    const dim3 grid(blocks_for(rows, kWG / G)), block(kWG);
    switch (epi) {
        case EPI_PLAIN: hipLaunchKernelGGL((csr_vector_kernel<G, EPI_PLAIN>), grid, block, 0, s, rowptr, col, val, x, rows, ea); break;
        case EPI_RATIO: hipLaunchKernelGGL((csr_vector_kernel<G, EPI_RATIO>), grid, block, 0, s, rowptr, col, val, x, rows, ea); break;
        default: hipLaunchKernelGGL((csr_vector_kernel<G, EPI_UPDATE>), grid, block, 0, s, rowptr, col, val, x, rows, ea); break;
    }
    return hipGetLastError();
}

template <int G>
hipError_t launch_scatter_g(const uint64_t *rowptr, const uint32_t *col, const float *val, const float *xrow,
                            float *out, uint32_t rows, hipStream_t s) {
    hipLaunchKernelGGL(scatter_kernel<G>, dim3(blocks_for(rows, kWG / G)), dim3(kWG), 0, s, rowptr, col, val, xrow,
                       out, rows);
    return hipGetLastError();
}

template <int G>
hipError_t launch_loglik_g(const uint64_t *rowptr, const uint32_t *col, const float *val, const float *x,
                           const int32_t *g, uint32_t rows, double *part, hipStream_t s) {
    hipLaunchKernelGGL(loglik_kernel<G>, dim3(blocks_for(rows, kWG / G)), dim3(kWG), 0, s, rowptr, col, val, x, g,
                       rows, part);
    return hipGetLastError();
}

constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

MergePlan merge_plan(uint32_t rows, uint64_t nnz, uint32_t secsize) {
    MergePlan p;
    p.rows = rows;
    p.nnz = nnz;
    p.secsize = secsize;
    p.tile_items = (uint64_t)kWG * secsize;
    const uint64_t L = (uint64_t)rows + nnz;
    p.ntiles = secsize ? (uint32_t)((L + p.tile_items - 1) / p.tile_items) : 0;
    return p;
}

size_t merge_partition_bytes(const MergePlan &p) {
    return align256(sizeof(uint32_t) * ((size_t)p.ntiles + 1)) + align256(sizeof(uint64_t) * ((size_t)p.ntiles + 1));
}

size_t merge_scratch_bytes(const MergePlan &p) { return 2 * align256(sizeof(float) * ((size_t)p.ntiles + 1)); }

hipError_t launch_merge_partition(const MergePlan &p, const uint64_t *rowptr, void *partition, hipStream_t s) {
    if (p.rows == 0) return hipSuccess;
    uint32_t *tile_row = static_cast<uint32_t *>(partition);
    uint64_t *tile_nz = reinterpret_cast<uint64_t *>(static_cast<char *>(partition) +
                                                     align256(sizeof(uint32_t) * ((size_t)p.ntiles + 1)));
    hipLaunchKernelGGL(merge_partition_kernel, dim3(blocks_for((uint64_t)p.ntiles + 1, kWG)), dim3(kWG), 0, s,
                       rowptr, p.rows, p.nnz, p.tile_items, p.ntiles, tile_row, tile_nz);
    return hipGetLastError();
}

hipError_t launch_merge_spmv(const MergePlan &p, const void *partition, void *scratch, const uint64_t *rowptr,
                             const uint32_t *col, const float *val, const float *x, Epi epi, const EpiArgs &ea,
                             hipStream_t s) {
    if (p.rows == 0) return hipSuccess;
    const uint32_t *tile_row = static_cast<const uint32_t *>(partition);
    const uint64_t *tile_nz = reinterpret_cast<const uint64_t *>(
        static_cast<const char *>(partition) + align256(sizeof(uint32_t) * ((size_t)p.ntiles + 1)));
    float *carry = static_cast<float *>(scratch);
    float *head = reinterpret_cast<float *>(static_cast<char *>(scratch) + align256(sizeof(float) * ((size_t)p.ntiles + 1)));
    switch (epi) {
        case EPI_PLAIN: return launch_merge_epi<EPI_PLAIN>(p, tile_row, tile_nz, carry, head, rowptr, col, val, x, ea, s);
        case EPI_RATIO: return launch_merge_epi<EPI_RATIO>(p, tile_row, tile_nz, carry, head, rowptr, col, val, x, ea, s);
        default: return launch_merge_epi<EPI_UPDATE>(p, tile_row, tile_nz, carry, head, rowptr, col, val, x, ea, s);
    }
}

bool valid_lanes(uint32_t lanes) { return lanes == 4 || lanes == 8 || lanes == 16 || lanes == 32 || lanes == 64; }

// Lanes per row from the mean row length: one 16-byte load per lane should
// cover a mean row in about one to two rounds.
uint32_t auto_lanes(uint32_t rows, uint64_t nnz) {
    const double m = rows ? (double)nnz / rows : 0.0;
    if (m <= 16) return 4;
    // Rows of a few dozen entries with 8 lanes take two rounds of loads
    // (one memory latency each); 16 lanes take one: on the 2D ring (48 per
    // row, at most 115) FP ran 15.0 -> 11.1 us per launch.
    if (m <= 32) return 8;
    if (m <= 128) return 16;
    if (m <= 384) return 32;
    return 64;
}

hipError_t launch_csr_vector(uint32_t lanes, const uint64_t *rowptr, const uint32_t *col, const float *val,
                             const float *x, uint32_t rows, Epi epi, const EpiArgs &ea, hipStream_t s) {
    if (rows == 0) return hipSuccess;
    switch (lanes) {
        case 4: return launch_csr_vector_g<4>(rowptr, col, val, x, rows, epi, ea, s);
        case 8: return launch_csr_vector_g<8>(rowptr, col, val, x, rows, epi, ea, s);
        case 16: return launch_csr_vector_g<16>(rowptr, col, val, x, rows, epi, ea, s);
        case 32: return launch_csr_vector_g<32>(rowptr, col, val, x, rows, epi, ea, s);
        default: return launch_csr_vector_g<64>(rowptr, col, val, x, rows, epi, ea, s);
    }
}

hipError_t launch_scatter(uint32_t lanes, const uint64_t *rowptr, const uint32_t *col, const float *val,
                          const float *xrow, float *out, uint32_t rows, hipStream_t s) {
    if (rows == 0) return hipSuccess;
    switch (lanes) {
        case 4: return launch_scatter_g<4>(rowptr, col, val, xrow, out, rows, s);
        case 8: return launch_scatter_g<8>(rowptr, col, val, xrow, out, rows, s);
        case 16: return launch_scatter_g<16>(rowptr, col, val, xrow, out, rows, s);
        case 32: return launch_scatter_g<32>(rowptr, col, val, xrow, out, rows, s);
        default: return launch_scatter_g<64>(rowptr, col, val, xrow, out, rows, s);
    }
}

hipError_t launch_ratio(const int32_t *g, float *fwproj, uint32_t rows, hipStream_t s) {
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(ratio_kernel, dim3(blocks_for(rows, kWG)), dim3(kWG), 0, s, g, fwproj, rows);
    return hipGetLastError();
}

hipError_t launch_update(const float *f, const float *norm, float *bwproj, uint32_t cols, hipStream_t s) {
    if (cols == 0) return hipSuccess;
    hipLaunchKernelGGL(update_kernel, dim3(blocks_for(cols, kWG)), dim3(kWG), 0, s, f, norm, bwproj, cols);
    return hipGetLastError();
}

hipError_t launch_update_image(float *f, const float *norm, float *bp, uint32_t cols, hipStream_t s) {
    if (cols == 0) return hipSuccess;
    hipLaunchKernelGGL(update_image_kernel, dim3(blocks_for(cols, kWG)), dim3(kWG), 0, s, f, norm, bp, cols);
    return hipGetLastError();
}

hipError_t launch_fill(float *dst, float v, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + kWG - 1) / kWG;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_kernel<float>, dim3((uint32_t)blocks), dim3(kWG), 0, s, dst, v, n);
    return hipGetLastError();
}

hipError_t launch_copy(float *dst, const float *src, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n / 4 + kWG - 1) / kWG;
    blocks = std::max<uint64_t>(1, std::min<uint64_t>(blocks, 8192));
    hipLaunchKernelGGL(copy_kernel, dim3((uint32_t)blocks), dim3(kWG), 0, s, dst, src, n);
    return hipGetLastError();
}

hipError_t launch_aos_split(const void *aos, uint64_t n, uint64_t first, uint32_t cols, uint32_t *col, float *val,
                            uint32_t *bad, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint64_t blocks = std::max<uint64_t>(1, std::min<uint64_t>((n + kWG - 1) / kWG, 16384));
    hipLaunchKernelGGL(aos_split_kernel, dim3((uint32_t)blocks), dim3(kWG), 0, s, static_cast<const uint32_t *>(aos),
                       n, first, cols, col, val, bad);
    return hipGetLastError();
}

hipError_t launch_first_bad_column(const uint32_t *col, uint64_t n, uint32_t cols, uint64_t *first, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint64_t blocks = std::max<uint64_t>(1, std::min<uint64_t>((n + kWG - 1) / kWG, 16384));
    hipLaunchKernelGGL(first_bad_column_kernel, dim3((uint32_t)blocks), dim3(kWG), 0, s, col, n, cols,
                       reinterpret_cast<unsigned long long *>(first));
    return hipGetLastError();
}

hipError_t launch_strided_put(void *dst, uint32_t stride, const void *src, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint64_t blocks = std::max<uint64_t>(1, std::min<uint64_t>((n + kWG - 1) / kWG, 8192));
    hipLaunchKernelGGL(strided_put_kernel<uint32_t>, dim3((uint32_t)blocks), dim3(kWG), 0, s,
                       static_cast<uint32_t *>(dst), stride, static_cast<const uint32_t *>(src), n);
    return hipGetLastError();
}

hipError_t launch_strided_get(void *dst, const void *src, uint32_t stride, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint64_t blocks = std::max<uint64_t>(1, std::min<uint64_t>((n + kWG - 1) / kWG, 8192));
    hipLaunchKernelGGL(strided_get_kernel<uint32_t>, dim3((uint32_t)blocks), dim3(kWG), 0, s,
                       static_cast<uint32_t *>(dst), static_cast<const uint32_t *>(src), stride, n);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *dst, uint32_t v, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + kWG - 1) / kWG;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3((uint32_t)blocks), dim3(kWG), 0, s, dst, v, n);
    return hipGetLastError();
}

hipError_t launch_fill_f64(double *dst, double v, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + kWG - 1) / kWG;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_kernel<double>, dim3((uint32_t)blocks), dim3(kWG), 0, s, dst, v, n);
    return hipGetLastError();
}

hipError_t launch_f64_to_f32(const double *a, float *b, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(f64_to_f32_kernel, dim3(blocks_for(n, kWG)), dim3(kWG), 0, s, a, b, n);
    return hipGetLastError();
}

hipError_t launch_colsum_rows_f64(const uint64_t *rowptrT, const float *valT, uint32_t cols, double *s64, float *s32,
                                  hipStream_t s) {
    if (cols == 0) return hipSuccess;
    hipLaunchKernelGGL(colsum_rows_kernel, dim3(blocks_for(cols, kWG / 64)), dim3(kWG), 0, s, rowptrT, valT, cols,
                       s64, s32);
    return hipGetLastError();
}

hipError_t launch_colsum_weighted_f64(const uint64_t *rowptrT, const uint32_t *rowT, const float *valT, const float *w,
                                      uint32_t cols, double *s64, float *s32, hipStream_t s) {
    if (cols == 0) return hipSuccess;
    hipLaunchKernelGGL(colsum_weighted_kernel, dim3(blocks_for(cols, kWG / 64)), dim3(kWG), 0, s, rowptrT, rowT, valT,
                       w, cols, s64, s32);
    return hipGetLastError();
}

hipError_t launch_colsum_scatter_f64(const uint64_t *rowptr, const uint32_t *col, const float *val, uint32_t rows,
                                     double *s64, hipStream_t s) {
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(colsum_scatter_kernel, dim3(blocks_for(rows, kWG / 64)), dim3(kWG), 0, s, rowptr, col, val,
                       rows, s64);
    return hipGetLastError();
}

uint32_t loglik_blocks(uint32_t rows) { return blocks_for(rows, kWG / 16); }

// The same terms from y = A x given (16 rows per block, as loglik_kernel<16>).
__global__ void loglik_y_kernel(const float *__restrict__ y, const int32_t *__restrict__ g, uint32_t rows,
                                double *__restrict__ part) {
    __shared__ double s[16];
    const uint32_t row = blockIdx.x * 16u + threadIdx.x;
    double l = 0.0;
    if (row < rows) {
        const double yd = (double)y[row];
        l = (yd > 0.0) ? (double)g[row] * log(yd) - yd : -yd;
    }
    s[threadIdx.x] = l;
    __syncthreads();
    if (threadIdx.x == 0) {
        double acc = 0.0;
        for (int q = 0; q < 16; ++q) acc += s[q];
        part[blockIdx.x] = acc;
    }
}

hipError_t launch_loglik_y(const float *y, const int32_t *g, uint32_t rows, double *part, hipStream_t s) {
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(loglik_y_kernel, dim3(loglik_blocks(rows)), dim3(16), 0, s, y, g, rows, part);
    return hipGetLastError();
}

hipError_t launch_loglik(const uint64_t *rowptr, const uint32_t *col, const float *val, const float *x,
                         const int32_t *g, uint32_t rows, double *part, hipStream_t s) {
    if (rows == 0) return hipSuccess;
    return launch_loglik_g<16>(rowptr, col, val, x, g, rows, part, s);
}

}  // namespace mlem
