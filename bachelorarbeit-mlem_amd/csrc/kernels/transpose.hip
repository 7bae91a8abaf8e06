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
#include <rocprim/device/device_radix_sort.hpp>

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

}  // namespace mlem
