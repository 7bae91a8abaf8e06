// Row orders for the batched layouts: the matrix with its rows (LORs) in a
// given order, built on the device from the resident CSR, and the scatter of
// a sinogram into that order.  With the rows of A grouped by the transaxial
// line of their LOR, the LORs through one voxel fall into few segments of
// A^T's columns, so A^T's quad tiles need less padding.
#include <rocprim/device/device_scan.hpp>

#include <algorithm>

#include "launch.hpp"

namespace mlem {
namespace {

// inv[order[i]] = i: the row of A that takes place j.
__global__ void invert_order_kernel(const uint32_t *__restrict__ order, uint32_t rows, uint32_t *__restrict__ inv) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < rows; i += gridDim.x * blockDim.x) inv[order[i]] = i;
}

// len[j] = entries of row inv[j]; len[rows] = 0 (the scan's last element).
__global__ void placed_lengths_kernel(const uint64_t *__restrict__ rowptr, const uint32_t *__restrict__ inv,
                                      uint32_t rows, uint64_t *__restrict__ len) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j <= rows; j += gridDim.x * blockDim.x) {
        if (j == rows) {
            len[j] = 0;
        } else {
            const uint32_t r = inv[j];
            len[j] = rowptr[r + 1] - rowptr[r];
        }
    }
}

// Row j of the reordered matrix is row inv[j] of A: one workgroup per row,
// consecutive work-items on consecutive entries.
__global__ __launch_bounds__(kWG) void place_rows_kernel(const uint64_t *__restrict__ rowptr,
                                                         const uint32_t *__restrict__ col,
                                                         const float *__restrict__ val,
                                                         const uint32_t *__restrict__ inv,
                                                         const uint64_t *__restrict__ rowptr_out,
                                                         uint32_t *__restrict__ col_out, float *__restrict__ val_out) {
    const uint32_t j = blockIdx.x;
    const uint32_t r = inv[j];
    const uint64_t src = rowptr[r], n = rowptr[r + 1] - src, dst = rowptr_out[j];
    for (uint64_t k = threadIdx.x; k < n; k += kWG) {
        col_out[dst + k] = col[src + k];
        val_out[dst + k] = val[src + k];
    }
}

// One slot of an interleaved [n][stride] array, in row order: dst[order[i] *
// stride] = src[i].
__global__ void ordered_put_kernel(uint32_t *__restrict__ dst, uint32_t stride, const uint32_t *__restrict__ src,
                                   const uint32_t *__restrict__ order, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        dst[(uint64_t)order[i] * stride] = src[i];
}

uint32_t grid_for(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + kWG - 1) / kWG, 8192)); }

}  // namespace

hipError_t reorder_rows(uint32_t rows, uint64_t nnz, const uint64_t *rowptr, const uint32_t *col, const float *val,
                        const uint32_t *order, uint64_t *rowptr_out, uint32_t *col_out, float *val_out,
                        hipStream_t s) {
    hipError_t e = hipSuccess;
    uint32_t *inv = nullptr;
    uint64_t *len = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    if ((e = hipMalloc(&inv, sizeof(uint32_t) * std::max<uint32_t>(1, rows))) != hipSuccess) return e;
    if ((e = hipMalloc(&len, sizeof(uint64_t) * ((size_t)rows + 1))) != hipSuccess) goto done;
    hipLaunchKernelGGL(invert_order_kernel, dim3(grid_for(rows)), dim3(kWG), 0, s, order, rows, inv);
    hipLaunchKernelGGL(placed_lengths_kernel, dim3(grid_for((uint64_t)rows + 1)), dim3(kWG), 0, s, rowptr, inv, rows,
                       len);
    if ((e = hipGetLastError()) != hipSuccess) goto done;
    if ((e = rocprim::exclusive_scan(nullptr, tmp_bytes, len, rowptr_out, uint64_t(0), (size_t)rows + 1,
                                     rocprim::plus<uint64_t>(), s)) != hipSuccess)
        goto done;
    if ((e = hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 4))) != hipSuccess) goto done;
    if ((e = rocprim::exclusive_scan(tmp, tmp_bytes, len, rowptr_out, uint64_t(0), (size_t)rows + 1,
                                     rocprim::plus<uint64_t>(), s)) != hipSuccess)
        goto done;
    if (rows && nnz)
        hipLaunchKernelGGL(place_rows_kernel, dim3(rows), dim3(kWG), 0, s, rowptr, col, val, inv, rowptr_out, col_out,
                           val_out);
    if ((e = hipGetLastError()) != hipSuccess) goto done;
    e = hipStreamSynchronize(s);
done:
    if (tmp) (void)hipFree(tmp);
    if (len) (void)hipFree(len);
    (void)hipFree(inv);
    return e;
}

hipError_t launch_ordered_put(void *dst, uint32_t stride, const void *src, const uint32_t *order, uint32_t n,
                              hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(ordered_put_kernel, dim3(grid_for(n)), dim3(kWG), 0, s, static_cast<uint32_t *>(dst), stride,
                       static_cast<const uint32_t *>(src), order, n);
    return hipGetLastError();
}

}  // namespace mlem
