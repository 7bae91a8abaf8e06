// Internal declarations shared by the C ABI's translation units (abi.hip,
// tiles_only.hip, batch.hip): error reporting, device buffers, the handles
// behind the opaque types of mlem_hip.h, and the helpers that build
// partitions, tiles and work lists on first use.  Not installed.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <array>
#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <tuple>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "launch.hpp"
#include "mlem_hip.h"
#include "mlem_host.h"

namespace mlem {
namespace abi {

// The calling thread's last error message (mlem_last_error()).
extern thread_local std::string g_err;
// Sets the message, returns st.
int set_err(int st, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
// A HIP error as a status (out of memory, no device, else a HIP error) with
// the failing call and its place.
int hip_err(hipError_t e, const char *what, const char *file, int line);

inline hipStream_t S(void *s) { return static_cast<hipStream_t>(s); }

}  // namespace abi
}  // namespace mlem

#define HIPCHK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) return mlem::abi::hip_err(e_, #expr, __FILE_NAME__, __LINE__); \
    } while (0)

#define ARGCHK(cond, ...)                                                                \
    do {                                                                                 \
        if (!(cond)) return mlem::abi::set_err(MLEM_ERR_INVALID_ARGUMENT, __VA_ARGS__);   \
    } while (0)

namespace mlem {
namespace abi {

// Device memory owned by a handle.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { reset(); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t n) {
        reset();
        bytes = n;
        if (n == 0) return hipSuccess;
        return hipMalloc(&p, n);
    }
    template <typename T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

// Host-side checks of a CSR's row pointers before it reaches the GPU (a bad
// one would make the kernels read outside their arrays); the column indices
// are checked on the device once uploaded.
int check_csr_host(const char *what, uint32_t rows, uint32_t cols, uint64_t nnz, const uint64_t *rowptr,
                   const uint32_t *col);
// Column indices of an uploaded CSR: MLEM_ERR_INVALID_ARGUMENT naming the
// first index >= cols.
int check_columns_device(const char *what, const uint32_t *col, uint64_t nnz, uint32_t cols);
// Host -> device copy, large ones through pinned buffers (waits).
int upload_h2d(void *dst, const void *src, size_t bytes);

inline uint32_t clamp_secsize(uint32_t s) { return std::min(std::max(s, 1u), kMaxSecsize); }

// The SpMV kernels read col/val with 16-byte vector loads.
inline bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

}  // namespace abi
}  // namespace mlem

// ---------------------------------------------------------------------------
// handles
// ---------------------------------------------------------------------------
struct Partition {
    mlem::MergePlan plan;
    mlem::abi::DevBuf buf;
};

// A work list of a tiled matrix on the device (one per unit size and
// target workgroup count).
struct TileWorkDev {
    mlem::TileWork view;
    mlem::abi::DevBuf mem;
};

struct Tiled {
    mlem::TilePlan plan;
    mlem::abi::DevBuf mem;
    mlem::TiledMatrix view;
    mlem::abi::DevBuf sens;  // A^T tiles of 16-bit values: the quantized matrix's column sums (f32)
    bool ordered = false;    // built with A's rows in the matrix's row order (mlem_matrix_set_row_order)
    std::map<uint32_t, std::unique_ptr<TileWorkDev>> works;  // by target workgroup count
};

struct mlem_matrix {
    int device = 0;
    uint32_t rows = 0, cols = 0;
    uint64_t nnz = 0;
    mlem::abi::DevBuf rowptr, col, val;
    mlem::abi::DevBuf rowptrT, colT, valT;
    bool has_T = false;
    // Tiles only (mlem_matrix_create_tiled): no CSR copies on the device, the
    // tiles of A and A^T (width tile_W, tile_R rows, f32) and the
    // sensitivity taken while they were built.
    bool tiles_only = false;
    uint32_t tile_W = 0, tile_R = 0, tile_vt = 0, tile_bc = 1;  // tile_bc: right-hand sides per pass
    mlem::abi::DevBuf sens32, sens64;
    // Row order for the batched (quad) tiles: row_order[i] = place of row i
    // of A (mlem_matrix_set_row_order); has_order false: A's own order.
    mlem::abi::DevBuf row_order;
    bool has_order = false;
    std::mutex mu;
    std::map<std::pair<int, uint32_t>, std::unique_ptr<Partition>> parts;  // (0: A / 1: A^T, secsize)
    // Tiles keyed by (0: A / 1: A^T, width W, rows R, value type, quad groups,
    // rows in the row order).
    using TileKey = std::tuple<int, uint32_t, uint32_t, uint32_t, bool, bool>;
    std::map<TileKey, std::unique_ptr<Tiled>> tiles;
    std::map<TileKey, uint32_t> tile_refused;  // build flags of refused tilings
    // Reconstructions and batches alive on this matrix: they hold its arrays,
    // tiles and work lists by address, so the transpose is not rebuilt and
    // the matrix is not destroyed while any exists.
    mutable std::atomic<uint32_t> users{0};

    uint64_t device_bytes() const {
        uint64_t b = rowptr.bytes + col.bytes + val.bytes + rowptrT.bytes + colT.bytes + valT.bytes + sens32.bytes +
                     sens64.bytes + row_order.bytes;
        for (const auto &t : tiles) b += t.second->mem.bytes;
        return b;
    }
};

// Per profiled iteration: FP start, FP SpMV kernel done, FP done, BP SpMV
// kernel done, BP done.
struct IterEvents {
    hipEvent_t e[5];
};

// One pass of a batch: the counts, ratios and images of its BC slots,
// [rows_x][BC] / [cols_x][BC] (right-hand side fastest).  Allocated when a
// slot of the pass is first set, so a batch holds device memory for the
// passes it has used, not for all its slots.
struct BatchPass {
    mlem::abi::DevBuf g, ratio, f;
};

struct mlem_batch {
    mlem_matrix *m = nullptr;
    uint32_t slots = 0, BC = 4, passes = 0;
    uint32_t nt = 0;                // work-items per workgroup of the stream kernel (0: its default)
    size_t rows_x = 0, cols_x = 0;  // padded vector lengths
    hipStream_t stream = nullptr;
    std::vector<std::unique_ptr<BatchPass>> pass;  // [passes], null until a slot of the pass is set
    // sensitivity, [slots] active flags, a rows/cols-long staging vector, partial sums of FP and BP
    mlem::abi::DevBuf sens, active, stage, part_fp, part_bp;
    const Tiled *tfp = nullptr, *tbp = nullptr;
    const TileWorkDev *wfp = nullptr, *wbp = nullptr;
    const uint32_t *row_order = nullptr;  // tiles in the matrix's row order: where each count of g goes
    std::vector<uint32_t> active_host;
    // One iteration captured as a graph (use_graph), for the passes that
    // held an active slot when it was captured.
    bool use_graph = false;
    hipGraphExec_t gexec = nullptr;
    std::vector<uint8_t> graph_passes;
    // Device bytes of one pass, and of everything else the batch holds.
    uint64_t pass_bytes() const { return (uint64_t)BC * 4u * (2u * rows_x + cols_x); }
    uint64_t fixed_bytes() const {
        return sens.bytes + active.bytes + stage.bytes + part_fp.bytes + part_bp.bytes;
    }
    uint64_t device_bytes() const {
        uint64_t b = fixed_bytes();
        for (const auto &q : pass)
            if (q) b += q->g.bytes + q->ratio.bytes + q->f.bytes;
        return b;
    }
    ~mlem_batch() {
        if (stream) (void)hipStreamSynchronize(stream);  // no replay still running
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

struct mlem_recon {
    const mlem_matrix *m = nullptr;
    mlem_opts opts{};
    hipStream_t stream = nullptr;
    mlem::abi::DevBuf g, sens, f, ratio, bp, scratch_fp, scratch_bp, llpart;
    const Partition *pfp = nullptr, *pbp = nullptr;
    const Tiled *tfp = nullptr, *tbp = nullptr;        // blocked LDS-tiled path, when used
    const TileWorkDev *wfp = nullptr, *wbp = nullptr;  // and its work lists
    mlem::abi::DevBuf part_fp, part_bp;
    uint32_t lanes_fp = 64, lanes_bp = 64;
    bool bp_merge = true;  // BP on A^T: merge-based or CSR-vector
    std::vector<IterEvents> events;  // pool, one set per profiled iteration
    size_t events_used = 0;
    double phase_ms[4] = {0, 0, 0, 0};  // FP SpMV, FP epilogue, BP SpMV, BP epilogue
    uint32_t timed = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
};

namespace mlem {
namespace abi {

// Merge-path partition of A (which 0) or A^T (1) for secsize, built on first use.
int get_partition(mlem_matrix *m, int which, uint32_t secsize, const Partition **out);
// Tiles of A or A^T, built on first use from the resident CSR (a
// tiles-only matrix has only the ones it was built with).
int get_tiled(mlem_matrix *m, int which, uint32_t W, uint32_t R, uint32_t vt, bool quad, const Tiled **out,
              uint32_t *flags);
// Work list of t for about `target` workgroups, built on first use.
int get_tile_work(mlem_matrix *m, Tiled *t, uint32_t target, const TileWorkDev **out);
// Tiles, work list and partial-sum buffer of one side for BC right-hand sides.
int setup_tiled(mlem_matrix *m, int which, uint32_t W, uint32_t R, uint32_t vt, uint32_t BC, uint32_t groups,
                const Tiled **tiled, const TileWorkDev **work, DevBuf *part);
// The blocked or the plain merge path for one side of a reconstruction.
int setup_merge_side(mlem_matrix *m, int which, const mlem_opts &o, uint32_t secsize, const Tiled **tiled,
                     const TileWorkDev **work, DevBuf *part, const Partition **partn, DevBuf *scratch);
// FP64 column sums of A (s32 and s64 may be null).
int sens_on_device(const mlem_matrix *m, float *s32, double *s64_or_null, hipStream_t s);
// Rows per tile block of the blocked kernels for BC right-hand sides when
// none is asked for, on the current device: 512 when the matrix's runs
// (about `groups` of them; 0: 24 per CU for one right-hand side, 8 for
// several) would hold under two units of entries each -- a run's partial
// sums (4 bytes per row of its block and right-hand side) would then cost
// about as much as its entries -- else 2048 (one) or 1024 (several).
uint32_t default_tile_rows(uint64_t nnz, uint32_t groups, uint32_t BC);
// Batch slots without waiting on the host (the serving loop): the set
// copies g (rows) and f0 (cols) in and marks the slot active, allocating its
// pass if that is the first slot set in it (*allocated = bytes allocated,
// may be null); the image fetch copies the slot's image out; deactivation
// clears the flag.  All are queued on the batch's stream: host arrays must
// stay valid, and the image is not there, until the stream has passed them.
int batch_set_async(mlem_batch *b, uint32_t slot, const int32_t *g, const float *f0, uint64_t *allocated);
int batch_get_image_async(mlem_batch *b, uint32_t slot, float *f_out);
int batch_deactivate_async(mlem_batch *b, uint32_t slot);
// The reconstruction in active slot `from` continues in free slot `to` (of a
// pass that holds memory): its counts and image are copied on the batch's
// stream, the flags swapped.  A slot's arithmetic does not depend on its
// place, so the images are bitwise those it would have had in `from`.
int batch_move_slot_async(mlem_batch *b, uint32_t from, uint32_t to);
// Frees pass p when none of its slots is active (returns the bytes freed).
uint64_t batch_release_pass(mlem_batch *b, uint32_t p);
// Progress of long builds on stderr when MLEM_VERBOSE is set.
void progress(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

}  // namespace abi
}  // namespace mlem
