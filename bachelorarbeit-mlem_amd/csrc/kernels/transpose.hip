// Device-side CSR(A) -> CSR(A^T): the stable counting sort of the nonzeros by
// column that sptrans_scanTrans_specialized (sptrans.h:397-545) computes on
// the CPU.  Stability is what makes the two agree bit for bit: a stable sort
// by column of the entries in nnz order keeps, inside every column, the
// original nnz (= row) order, which is exactly sptrans' placement.
//
//   1. pack   (row, value bits) of every nonzero into one u64, in nnz order
//   2. stable LSD radix sort of (col, packed) pairs on the ceil(log2 cols)
//      low key bits (rocPRIM radix_sort_pairs)
//   3. unpack rowT / valT; colptrT[c] = lower_bound(sorted cols, c)
//
// That needs 20 bytes of scratch per nonzero.  Where it does not fit (the
// 100 GB matrices), the same sort runs over column ranges one at a time:
// colptrT from a column histogram and a scan; then for every range
// [c0, c1) the nonzeros with a column in it are compacted in nnz order
// (a wave per row, ballot ranks), sorted stably by column and unpacked at
// colptrT[c0].  Every range is the slice of the full sort's output for
// those columns, so the result is the same bit for bit.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <vector>

#include "launch.hpp"

namespace mlem {
namespace {

__global__ __launch_bounds__(kWG) void pack_rows_kernel(const uint64_t *__restrict__ rowptr,
                                                        const float *__restrict__ val, uint32_t rows,
                                                        uint64_t *__restrict__ packed) {
    const uint32_t row = blockIdx.x * (kWG / 64) + threadIdx.x / 64;
    const uint32_t lane = threadIdx.x % 64;
    if (row >= rows) return;
    const uint64_t hi = (uint64_t)row << 32;
    for (uint64_t k = rowptr[row] + lane; k < rowptr[row + 1]; k += 64)
        packed[k] = hi | (uint64_t)__float_as_uint(val ? val[k] : 0.0f);
}

__global__ void unpack_kernel(const uint64_t *__restrict__ packed, uint64_t nnz, uint32_t *__restrict__ rowT,
                              float *__restrict__ valT) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < nnz; k += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t p = packed[k];
        rowT[k] = (uint32_t)(p >> 32);
        if (valT) valT[k] = __uint_as_float((uint32_t)p);
    }
}

__global__ void colptr_kernel(const uint32_t *__restrict__ keys, uint64_t nnz, uint32_t cols,
                              uint64_t *__restrict__ colptrT) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > cols) return;
    uint64_t lo = 0, hi = nnz;  // first position with keys[pos] >= c
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (keys[mid] < c)
            lo = mid + 1;
        else
            hi = mid;
    }
    colptrT[c] = lo;
}

__global__ void col_hist_kernel(const uint32_t *__restrict__ col, uint64_t nnz, uint32_t *__restrict__ cnt) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < nnz; k += (uint64_t)gridDim.x * blockDim.x)
        atomicAdd(cnt + col[k], 1u);
}

// cnt[row] = entries of the row with c0 <= column < c1.
__global__ __launch_bounds__(kWG) void range_count_kernel(const uint64_t *__restrict__ rowptr,
                                                          const uint32_t *__restrict__ col, uint32_t rows,
                                                          uint32_t c0, uint32_t c1, uint64_t *__restrict__ cnt) {
    const uint32_t row = blockIdx.x * (kWG / 64) + threadIdx.x / 64;
    const uint32_t lane = threadIdx.x % 64;
    if (row >= rows) return;
    uint32_t n = 0;
    for (uint64_t k = rowptr[row] + lane; k < rowptr[row + 1]; k += 64) n += (col[k] >= c0 && col[k] < c1) ? 1u : 0u;
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o);
    if (lane == 0) cnt[row] = n;
}

// The row's entries with a column in [c0, c1), in order, from base[row] on:
// key = column - c0, packed = row << 32 | value bits.
__global__ __launch_bounds__(kWG) void range_compact_kernel(const uint64_t *__restrict__ rowptr,
                                                            const uint32_t *__restrict__ col,
                                                            const float *__restrict__ val, uint32_t rows, uint32_t c0,
                                                            uint32_t c1, const uint64_t *__restrict__ base,
                                                            uint32_t *__restrict__ keys,
                                                            uint64_t *__restrict__ packed) {
    const uint32_t row = blockIdx.x * (kWG / 64) + threadIdx.x / 64;
    const uint32_t lane = threadIdx.x % 64;
    if (row >= rows) return;
    const uint64_t hi = (uint64_t)row << 32;
    uint64_t out = base[row];
    const uint64_t e = rowptr[row + 1];
    for (uint64_t k0 = rowptr[row]; k0 < e; k0 += 64) {
        const uint64_t k = k0 + lane;
        const uint32_t c = k < e ? col[k] : 0u;
        const bool in = k < e && c >= c0 && c < c1;
        const uint64_t m = __ballot(in);
        if (in) {
            const uint64_t at = out + __popcll(m & ((1ull << lane) - 1ull));
            keys[at] = c - c0;
            packed[at] = hi | (uint64_t)__float_as_uint(val ? val[k] : 0.0f);
        }
        out += __popcll(m);
    }
}

inline unsigned key_bits(uint32_t cols) {
    unsigned b = 1;
    while (b < 32 && ((uint64_t)1 << b) < cols) ++b;
    return b;
}

constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

size_t sort_temp_bytes(uint32_t cols, uint64_t nnz) {
    size_t bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                              (const uint64_t *)nullptr, (uint64_t *)nullptr, (size_t)nnz, 0u, key_bits(cols));
    return bytes;
}

// Scratch of one column range of up to n nonzeros: keys and packed, in and
// out, the sort's temporary storage.
size_t range_scratch_bytes(uint32_t cols, uint64_t n) {
    return 2 * align256(n * 4) + 2 * align256(n * 8) + align256(sort_temp_bytes(cols, n));
}

}  // namespace

size_t transpose_scratch_bytes(uint32_t rows, uint32_t cols, uint64_t nnz) {
    (void)rows;
    return align256(nnz * 8) * 2 + align256(nnz * 4) + align256(sort_temp_bytes(cols, nnz));
}

hipError_t launch_transpose(uint32_t rows, uint32_t cols, uint64_t nnz, const uint64_t *rowptr, const uint32_t *col,
                            const float *val, uint64_t *colptrT, uint32_t *rowT, float *valT, void *scratch,
                            hipStream_t s) {
    char *p = static_cast<char *>(scratch);
    uint64_t *packed_in = reinterpret_cast<uint64_t *>(p);
    p += align256(nnz * 8);
    uint64_t *packed_out = reinterpret_cast<uint64_t *>(p);
    p += align256(nnz * 8);
    uint32_t *keys_out = reinterpret_cast<uint32_t *>(p);
    p += align256(nnz * 4);
    size_t temp_bytes = sort_temp_bytes(cols, nnz);
    void *temp = p;
    if (nnz > 0) {
        hipLaunchKernelGGL(pack_rows_kernel, dim3((rows + kWG / 64 - 1) / (kWG / 64)), dim3(kWG), 0, s, rowptr, val,
                           rows, packed_in);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        e = rocprim::radix_sort_pairs(temp, temp_bytes, col, keys_out, packed_in, packed_out, (size_t)nnz, 0u,
                                      key_bits(cols), s);
        if (e != hipSuccess) return e;
        uint64_t blocks = (nnz + kWG - 1) / kWG;
        if (blocks > 65536) blocks = 65536;
        hipLaunchKernelGGL(unpack_kernel, dim3((uint32_t)blocks), dim3(kWG), 0, s, packed_out, nnz, rowT, valT);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(colptr_kernel, dim3((cols + 1 + kWG - 1) / kWG), dim3(kWG), 0, s, keys_out, nnz, cols,
                       colptrT);
    return hipGetLastError();
}

hipError_t transpose_in_ranges(uint32_t rows, uint32_t cols, uint64_t nnz, const uint64_t *rowptr,
                               const uint32_t *col, const float *val, uint64_t *colptrT, uint32_t *rowT, float *valT,
                               size_t budget, uint32_t *ranges_out, hipStream_t s) {
    return transpose_ranges_to(rows, cols, nnz, rowptr, col, val, colptrT, rowT, valT, budget, 1, RangeSink(),
                               ranges_out, s);
}

hipError_t transpose_ranges_to(uint32_t rows, uint32_t cols, uint64_t nnz, const uint64_t *rowptr,
                               const uint32_t *col, const float *val, uint64_t *colptrT, uint32_t *rowT, float *valT,
                               size_t budget, uint32_t align, const RangeSink &sink, uint32_t *ranges_out,
                               hipStream_t s) {
    hipError_t e;
    *ranges_out = 0;
    if (align == 0) align = 1;
    // colptrT: column counts, scanned (u32 counts, u64 offsets).
    size_t scan_bytes = 0;
    (void)rocprim::exclusive_scan(nullptr, scan_bytes, (const uint32_t *)nullptr, (uint64_t *)nullptr, uint64_t(0),
                                  (size_t)cols + 1, rocprim::plus<uint64_t>());
    size_t scan_rows = 0;
    (void)rocprim::exclusive_scan(nullptr, scan_rows, (const uint64_t *)nullptr, (uint64_t *)nullptr, uint64_t(0),
                                  (size_t)rows + 1, rocprim::plus<uint64_t>());
    const size_t fixed = align256(4 * ((size_t)cols + 1)) + align256(std::max(scan_bytes, scan_rows)) +
                         2 * align256(8 * ((size_t)rows + 1));
    if (budget <= fixed + range_scratch_bytes(cols, 1)) return hipErrorOutOfMemory;
    // The largest range that fits the rest of the budget.
    uint64_t lo = 1, hi = std::max<uint64_t>(nnz, 1);
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo + 1) / 2;
        if (fixed + range_scratch_bytes(cols, mid) <= budget)
            lo = mid;
        else
            hi = mid - 1;
    }
    const uint64_t max_n = lo;
    void *mem = nullptr;
    if ((e = hipMalloc(&mem, fixed + range_scratch_bytes(cols, max_n))) != hipSuccess) return e;
    char *p = static_cast<char *>(mem);
    uint32_t *ccnt = reinterpret_cast<uint32_t *>(p);
    p += align256(4 * ((size_t)cols + 1));
    void *scan_tmp = p;
    p += align256(std::max(scan_bytes, scan_rows));
    uint64_t *rcnt = reinterpret_cast<uint64_t *>(p);
    p += align256(8 * ((size_t)rows + 1));
    uint64_t *rbase = reinterpret_cast<uint64_t *>(p);
    p += align256(8 * ((size_t)rows + 1));
    uint32_t *keys_in = reinterpret_cast<uint32_t *>(p);
    p += align256(max_n * 4);
    uint32_t *keys_out = reinterpret_cast<uint32_t *>(p);
    p += align256(max_n * 4);
    uint64_t *packed_in = reinterpret_cast<uint64_t *>(p);
    p += align256(max_n * 8);
    uint64_t *packed_out = reinterpret_cast<uint64_t *>(p);
    p += align256(max_n * 8);
    void *sort_tmp = p;
    auto done = [&](hipError_t r) {
        (void)hipStreamSynchronize(s);
        (void)hipFree(mem);
        return r;
    };
    const dim3 row_grid((rows + kWG / 64 - 1) / (kWG / 64));
    if ((e = hipMemsetAsync(ccnt, 0, 4 * ((size_t)cols + 1), s)) != hipSuccess) return done(e);
    if (nnz) {
        uint64_t blocks = std::min<uint64_t>((nnz + kWG - 1) / kWG, 65536);
        hipLaunchKernelGGL(col_hist_kernel, dim3((uint32_t)blocks), dim3(kWG), 0, s, col, nnz, ccnt);
        if ((e = hipGetLastError()) != hipSuccess) return done(e);
    }
    size_t sb = scan_bytes;
    if ((e = rocprim::exclusive_scan(scan_tmp, sb, ccnt, colptrT, uint64_t(0), (size_t)cols + 1,
                                     rocprim::plus<uint64_t>(), s)) != hipSuccess)
        return done(e);
    std::vector<uint64_t> cp((size_t)cols + 1);
    if ((e = hipMemcpyAsync(cp.data(), colptrT, 8 * ((size_t)cols + 1), hipMemcpyDeviceToHost, s)) != hipSuccess)
        return done(e);
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return done(e);
    for (uint32_t c0 = 0; c0 < cols;) {
        // At least `align` columns (they must fit), else the furthest c1 with
        // cp[c1] - cp[c0] <= max_n, a multiple of align or cols.
        uint32_t c1 = (uint32_t)std::min<uint64_t>((uint64_t)c0 + align, cols);
        if (cp[c1] - cp[c0] > max_n) return done(hipErrorOutOfMemory);
        {
            uint32_t a = c1, b = cols;
            while (a < b) {
                const uint32_t mid = a + (b - a + 1) / 2;
                if (cp[mid] - cp[c0] <= max_n)
                    a = mid;
                else
                    b = mid - 1;
            }
            c1 = a == cols ? a : std::max(c1, a / align * align);
        }
        const uint64_t n = cp[c1] - cp[c0];
        ++*ranges_out;
        if (n) {
            hipLaunchKernelGGL(range_count_kernel, row_grid, dim3(kWG), 0, s, rowptr, col, rows, c0, c1, rcnt);
            if ((e = hipMemsetAsync(rcnt + rows, 0, 8, s)) != hipSuccess) return done(e);
            size_t rb = scan_rows;
            if ((e = rocprim::exclusive_scan(scan_tmp, rb, rcnt, rbase, uint64_t(0), (size_t)rows + 1,
                                             rocprim::plus<uint64_t>(), s)) != hipSuccess)
                return done(e);
            hipLaunchKernelGGL(range_compact_kernel, row_grid, dim3(kWG), 0, s, rowptr, col, val, rows, c0, c1, rbase,
                               keys_in, packed_in);
            if ((e = hipGetLastError()) != hipSuccess) return done(e);
            size_t tb = sort_temp_bytes(cols, max_n);
            if ((e = rocprim::radix_sort_pairs(sort_tmp, tb, keys_in, keys_out, packed_in, packed_out, (size_t)n, 0u,
                                               key_bits(c1 - c0), s)) != hipSuccess)
                return done(e);
            const uint64_t blocks = std::min<uint64_t>((n + kWG - 1) / kWG, 65536);
            // With a sink the range goes to the freed sort inputs (a slab of
            // A^T's rows c0..c1), else to its place in rowT / valT.
            uint32_t *rT = sink ? keys_in : rowT + cp[c0];
            float *vT = sink ? reinterpret_cast<float *>(packed_in) : (valT ? valT + cp[c0] : nullptr);
            hipLaunchKernelGGL(unpack_kernel, dim3((uint32_t)blocks), dim3(kWG), 0, s, packed_out, n, rT, vT);
            if ((e = hipGetLastError()) != hipSuccess) return done(e);
        }
        if (sink && (e = sink(c0, c1, cp[c0], n, keys_in, reinterpret_cast<const float *>(packed_in))) != hipSuccess)
            return done(e);
        c0 = c1;
    }
    return done(hipSuccess);
}

__global__ void rebase_kernel(const uint64_t *__restrict__ src, uint64_t base, uint32_t n, uint64_t *__restrict__ dst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) dst[i] = src[i] - base;
}

hipError_t launch_rebase(const uint64_t *src, uint64_t base, uint32_t n, uint64_t *dst, hipStream_t s) {
    hipLaunchKernelGGL(rebase_kernel, dim3(n / kWG + 1), dim3(kWG), 0, s, src, base, n, dst);
    return hipGetLastError();
}

}  // namespace mlem
