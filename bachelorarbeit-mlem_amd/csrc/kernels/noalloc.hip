// mlem_fp_merge without a workspace: for callers that cannot allocate device
// memory (code under stream capture, tight serving loops, harnesses that
// forbid hipMalloc).  The merge path's tile partition and cross-tile carries
// live in a fixed device array of this file instead of a caller-provided
// workspace; each device holds its own copy of the array, located once per
// device with hipGetSymbolAddress.  Calls that overlap on one device share
// that array and race: issue them on one stream, or wait between them.
#include "abi_internal.hpp"

using namespace mlem;
using namespace mlem::abi;

namespace {

// Scratch of the no-allocation path: partition (u32 + u64 per merge tile)
// plus carries and heads (2 x f32 per tile), about 20 bytes per tile of
// 256 * secsize merge items -- 1.6 million tiles, e.g. rows + nnz up to
// 3.4e9 at secsize 8.  Every byte read is written by the same call first.
static __device__ __align__(256) uint8_t scratch[MLEM_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The array's device address on the current device (hipGetSymbolAddress on
// first use there).  The array's name is a host-side stand-in: only this
// address may reach kernels and copies.
int scratch_address(void **out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    ARGCHK(dev >= 0 && dev < kMaxDevices, "mlem_fp_merge_noalloc: device %d beyond the %d supported", dev,
           kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ARGCHK(p != nullptr, "mlem_fp_merge_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return MLEM_OK;
}

}  // namespace

extern "C" {

int mlem_fp_merge_noalloc(const uint64_t *rowptr, const uint32_t *col, const float *val, const float *f,
                          float *fwproj, uint32_t secsize, uint32_t rows, uint64_t nnz, void *stream) {
    ARGCHK(secsize >= 1 && secsize <= kMaxSecsize, "mlem_fp_merge_noalloc: secsize %u outside 1..%u", secsize,
           kMaxSecsize);
    if (rows == 0) return MLEM_OK;
    ARGCHK(rowptr && f && fwproj && (nnz == 0 || (col && val)), "mlem_fp_merge_noalloc: null pointer");
    ARGCHK(aligned16(col) && aligned16(val), "mlem_fp_merge_noalloc: col and val must be 16-byte aligned");
    const MergePlan p = merge_plan(rows, nnz, secsize);
    const size_t need = merge_partition_bytes(p) + merge_scratch_bytes(p);
    ARGCHK(need <= (size_t)MLEM_NOALLOC_SCRATCH_BYTES,
           "mlem_fp_merge_noalloc: %zu bytes of scratch needed (rows + nnz = %llu at secsize %u), the fixed scratch "
           "holds %zu: use a larger secsize, or mlem_fp_merge with a workspace",
           need, (unsigned long long)((uint64_t)rows + nnz), secsize, (size_t)MLEM_NOALLOC_SCRATCH_BYTES);
    void *ws = nullptr;
    if (int st = scratch_address(&ws)) return st;
    EpiArgs ea{};
    ea.out = fwproj;
    HIPCHK(launch_merge_partition(p, rowptr, ws, S(stream)));
    HIPCHK(launch_merge_spmv(p, ws, static_cast<char *>(ws) + merge_partition_bytes(p), rowptr, col, val, f,
                             EPI_PLAIN, ea, S(stream)));
    return MLEM_OK;
}

}  // extern "C"
