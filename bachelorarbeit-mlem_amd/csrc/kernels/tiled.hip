// Blocked SpMV with x segments staged in LDS: the high-bandwidth variant of
// the merge-based kernels (calcFwProj_merge_based / calcBwProj_merge_based,
// kernel.cu:52-105,190-286).
//
// Why: a gather x[col] that misses the 32 KiB L1 costs a 128-byte line fill,
// and the L1 fill rate (about 64 bytes per clock per CU) -- not HBM -- caps
// CSR kernels that gather from global memory near 2.6-3 TB/s on the 12 GB
// scanner matrix, while streaming the same arrays runs at 6-7 TB/s.  Here the
// matrix is cut into tiles of R rows x W columns; a tile is dense enough
// (thousands of entries) that copying its W-wide segment of x into LDS once
// costs little, and every gather then reads LDS.
//
// Layout, built on the device from a CSR with column-sorted rows:
//   tile t = (row block b, segment s), t = b*nseg + s, stored in (b, s)
//   order, each tile padded to a multiple of 8 entries; inside a tile the
//   entries sorted by (row, column), so every row still sums its entries in
//   the original column order.  Per entry: value f32, column - s*W (u16),
//   row - b*R (u16): 8 bytes, the same as CSR's f32 value + u32 column.
//   tile_off u64[ntiles+1] (padded offsets), tile_nz u32[ntiles].
//
// Work: a tile is cut into units of 256*E entries; work-item k owns the E
// consecutive entries [kE, kE+E) of a unit (16-byte loads).  The units of a
// row block (its non-empty tiles in segment order) are split into runs of
// nearly equal length, one run per workgroup, so every workgroup streams
// about the same number of bytes whatever the tiles' fill (a work list built
// on the host from the tile sizes).  Row ends are where the row key changes.
// A row split between work-items is finished by the work-item holding its
// last entry, which adds the carries of the work-items before it (nearest
// first, back to the previous row end) from LDS; so each row that has
// entries in a unit is updated by exactly one work-item per unit, and the
// LDS accumulator of the block's rows needs no atomics (a read and a write,
// ordered by the unit's barriers).  The loads of the next unit (and, for a
// new tile, its x segment) are issued before the current unit is computed.
// Workgroup w writes its partials to part[w][0..R); a reduction kernel adds
// the workgroups of each row block in order and applies the epilogue (ratio
// or update).  Bitwise reproducible.
#include <rocprim/device/device_scan.hpp>

#include <algorithm>

#include "launch.hpp"

namespace mlem {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

template <typename T>
__device__ __forceinline__ T ldnt(const T *p) {
    return __builtin_nontemporal_load(p);
}

constexpr uint32_t kNoKey = 0x10000u;  // past every u16 row key

// A run position: staged tile ti (index into the workgroup's tile list),
// unit c of that tile, n entries in the unit.
struct Unit {
    uint32_t ti, c, n;
    bool valid;
};

template <int E>
__device__ __forceinline__ Unit make_unit(const uint4 *tl, uint32_t ti, uint32_t c, bool valid) {
    constexpr uint32_t UE = kWG * E;
    Unit u;
    u.ti = ti;
    u.c = c;
    u.valid = valid;
    const uint32_t tj = valid ? ti : 0u;  // ti < span whenever valid
    u.n = valid ? min(UE, tl[tj].y - c * UE) : 0u;
    return u;
}

// wl[w] = {row block, first tile (index into tiles), first unit in it, units}
// tiles[i] = {segment, entries, padded offset lo, hi} of the non-empty tiles.
template <int E, int W>
__global__ __launch_bounds__(kWG) void blocked_spmv_kernel(uint32_t R, uint32_t span, uint32_t ntl,
                                                           const uint4 *__restrict__ wl,
                                                           const uint4 *__restrict__ tiles,
                                                           const float *__restrict__ tval,
                                                           const uint16_t *__restrict__ tcol,
                                                           const uint16_t *__restrict__ trow,
                                                           const float *__restrict__ x, float *__restrict__ part) {
    constexpr uint32_t UE = kWG * E;
    constexpr int XR = W / kWG;
    extern __shared__ __attribute__((aligned(16))) float smem[];
    __shared__ uint32_t s_k0[kWG + 1];  // first key of every work-item (+ sentinel)
    __shared__ float s_cf[kWG];         // carry after the last row end
    __shared__ uint32_t s_hf[kWG];      // work-item saw a row end
    float *xs = smem;                                      // [W]
    float *acc = xs + W;                                   // [R]
    uint4 *tl = reinterpret_cast<uint4 *>(acc + R);        // [span]

    const uint32_t tid = threadIdx.x;
    const uint4 w = wl[blockIdx.x];
    for (uint32_t i = tid; i < R; i += kWG) acc[i] = 0.0f;
    for (uint32_t i = tid; i < span; i += kWG)
        if (w.y + i < ntl) tl[i] = tiles[w.y + i];
    if (tid == 0) s_k0[kWG] = kNoKey;
    __syncthreads();

    uint32_t left = w.w;  // units of the run not yet started
    Unit cur = make_unit<E>(tl, 0, w.z, left > 0);

    f32x4 pv[E / 4];
    u16x8 pc[E / 8], pr[E / 8];
    f32x4 px[XR / 4];
    // Loads of unit u; an invalid u re-reads the start of the current tile
    // (every address stays inside the arrays, so the wait counts are exact).
    auto prefetch = [&](const Unit &u, uint32_t fallback_ti) {
        const uint4 t = tl[u.valid ? u.ti : fallback_ti];
        const uint64_t off = ((uint64_t)t.w << 32) | t.z;
        uint64_t e = off + (uint64_t)(u.valid ? u.c : 0u) * UE + (uint64_t)tid * E;
        if ((uint64_t)tid * E >= (u.valid ? u.n : 1u)) e = off;  // stay inside the tile
#pragma unroll
        for (int m = 0; m < E / 4; ++m) pv[m] = ldnt(reinterpret_cast<const f32x4 *>(tval + e) + m);
#pragma unroll
        for (int m = 0; m < E / 8; ++m) {
            pc[m] = ldnt(reinterpret_cast<const u16x8 *>(tcol + e) + m);
            pr[m] = ldnt(reinterpret_cast<const u16x8 *>(trow + e) + m);
        }
        // x is padded to whole segments (xlen >= nseg * W): no clamping.
        const float *xb = x + (uint64_t)t.x * W;
#pragma unroll
        for (int k = 0; k < XR / 4; ++k) px[k] = *(reinterpret_cast<const f32x4 *>(xb) + tid + kWG * k);
    };

    if (cur.valid) prefetch(cur, 0);
    bool fresh = true;  // first unit of the run: its x segment is not staged yet
    while (cur.valid) {
        f32x4 cv[E / 4];
        u16x8 cc[E / 8], cr[E / 8];
        f32x4 cx[XR / 4];
#pragma unroll
        for (int m = 0; m < E / 4; ++m) cv[m] = pv[m];
#pragma unroll
        for (int m = 0; m < E / 8; ++m) {
            cc[m] = pc[m];
            cr[m] = pr[m];
        }
#pragma unroll
        for (int k = 0; k < XR / 4; ++k) cx[k] = px[k];
        --left;
        const bool more_in_tile = (cur.c + 1) * UE < tl[cur.ti].y;
        const Unit nxt = make_unit<E>(tl, more_in_tile ? cur.ti : cur.ti + 1, more_in_tile ? cur.c + 1 : 0u,
                                      left > 0);
        prefetch(nxt, cur.ti);  // in flight while this unit is computed

        // Keys of this work-item's entries; entries past n get kNoKey.
        const uint32_t e0 = tid * E;
        uint32_t key[E + 1];
#pragma unroll
        for (int m = 0; m < E; ++m) key[m] = (e0 + m < cur.n) ? (uint32_t)cr[m / 8][m % 8] : kNoKey;
        s_k0[tid] = key[0];
        if (cur.c == 0 || fresh) {  // a new tile: its x segment
#pragma unroll
            for (int k = 0; k < XR / 4; ++k) *(reinterpret_cast<f32x4 *>(xs) + tid + kWG * k) = cx[k];
        }
        fresh = false;
        __syncthreads();
        key[E] = s_k0[tid + 1];

        // Products first (all LDS gathers in flight together; padding and
        // masked entries read a valid column and are zeroed).
        float prod[E];
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const float p = cv[m / 4][m % 4] * xs[cc[m / 8][m % 8]];
            prod[m] = key[m] != kNoKey ? p : 0.0f;
        }
        // Walk my entries: a row ends where the key changes.  A row that
        // starts and ends here is complete: this work-item is its only
        // writer in this unit.
        float a = 0.0f, head = 0.0f;
        uint32_t has_end = 0, first_key = 0;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            a += prod[m];
            const bool end = (key[m] != key[m + 1]) && (key[m] != kNoKey);
            if (end && has_end) acc[key[m]] += a;
            const bool first = end && !has_end;
            head = first ? a : head;
            first_key = first ? key[m] : first_key;
            has_end |= end ? 1u : 0u;
            a = end ? 0.0f : a;
        }
        s_cf[tid] = a;
        s_hf[tid] = has_end;
        __syncthreads();
        // The first row ending here may have started in earlier work-items:
        // add their carries, nearest first, back to the last one with a row
        // end (whose carry belongs to this row too).
        if (has_end) {
            float cin = 0.0f;
            for (int jj = (int)tid - 1; jj >= 0; --jj) {
                cin += s_cf[jj];
                if (s_hf[jj]) break;
            }
            acc[first_key] += cin + head;
        }
        cur = nxt;
    }
    __syncthreads();
    float *pw = part + (uint64_t)blockIdx.x * R;
    for (uint32_t i = tid; i < R; i += kWG) pw[i] = acc[i];
}

// Row r of block b = r / R sums the partials of the block's workgroups
// wl_first[b] .. wl_first[b+1]-1 in order, then applies the epilogue.
template <int EPI>
__global__ void blocked_reduce_kernel(const float *__restrict__ part, const uint32_t *__restrict__ wl_first,
                                      uint32_t rshift, uint32_t rows, EpiArgs ea) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const uint32_t b = r >> rshift, i = r & ((1u << rshift) - 1u);
    float y = 0.0f;
    for (uint32_t w = wl_first[b]; w < wl_first[b + 1]; ++w) y += part[((uint64_t)w << rshift) + i];
    if constexpr (EPI == EPI_PLAIN) {
        ea.out[r] = y;
    } else if constexpr (EPI == EPI_RATIO) {
        ea.out[r] = (y != 0.0f) ? (float)ea.g[r] / y : y;
    } else {
        const float v = ea.f[r] * y;
        const float n = ea.norm[r];
        ea.f[r] = (n == 0.0f) ? v : v / n;
    }
}

// ----------------------------------------------------------------- builder
// cnt[r*nseg + s] = entries of row r in segment s.
__global__ __launch_bounds__(kWG) void tile_count_kernel(const uint64_t *__restrict__ rowptr,
                                                         const uint32_t *__restrict__ col, uint32_t rows,
                                                         uint32_t nseg, uint32_t wshift,
                                                         uint32_t *__restrict__ cnt) {
    const uint32_t row = blockIdx.x * (kWG / 64) + threadIdx.x / 64;
    const uint32_t lane = threadIdx.x % 64;
    if (row >= rows) return;
    for (uint64_t k = rowptr[row] + lane; k < rowptr[row + 1]; k += 64)
        atomicAdd(cnt + (uint64_t)row * nseg + (col[k] >> wshift), 1u);
}

// Per tile (one work-item per tile): rowpos[r*nseg+s] = entries of tile
// (b, s) in the rows of block b before r; exact and padded tile sizes.
__global__ void tile_rowpos_kernel(const uint32_t *__restrict__ cnt, uint32_t rows, uint32_t R, uint32_t nseg,
                                   uint64_t ntiles, uint32_t *__restrict__ rowpos, uint32_t *__restrict__ tile_nz,
                                   uint64_t *__restrict__ tile_padded) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const uint32_t b = (uint32_t)(t / nseg), s = (uint32_t)(t % nseg);
    uint32_t run = 0;
    for (uint32_t i = 0; i < R; ++i) {
        const uint64_t r = (uint64_t)b * R + i;
        if (r >= rows) break;
        rowpos[r * nseg + s] = run;
        run += cnt[r * nseg + s];
    }
    tile_nz[t] = run;
    tile_padded[t] = (run + 7u) & ~7u;
}

// Per row: exclusive prefix of its segment counts (in place).
__global__ void row_segscan_kernel(uint32_t *__restrict__ cnt, uint32_t rows, uint32_t nseg) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    uint32_t *c = cnt + (uint64_t)r * nseg;
    uint32_t run = 0;
    for (uint32_t s = 0; s < nseg; ++s) {
        const uint32_t n = c[s];
        c[s] = run;
        run += n;
    }
}

__global__ __launch_bounds__(kWG) void tile_scatter_kernel(const uint64_t *__restrict__ rowptr,
                                                           const uint32_t *__restrict__ col,
                                                           const float *__restrict__ val, uint32_t rows, uint32_t R,
                                                           uint32_t nseg, uint32_t wshift,
                                                           const uint32_t *__restrict__ segoff,
                                                           const uint32_t *__restrict__ rowpos,
                                                           const uint64_t *__restrict__ tile_off,
                                                           float *__restrict__ tval, uint16_t *__restrict__ tcol,
                                                           uint16_t *__restrict__ trow) {
    const uint32_t row = blockIdx.x * (kWG / 64) + threadIdx.x / 64;
    const uint32_t lane = threadIdx.x % 64;
    if (row >= rows) return;
    const uint32_t b = row / R, i = row % R;
    const uint64_t rs = rowptr[row];
    for (uint64_t k = rs + lane; k < rowptr[row + 1]; k += 64) {
        const uint32_t c = col[k];
        const uint32_t s = c >> wshift;
        const uint64_t t = (uint64_t)b * nseg + s;
        const uint64_t rsn = (uint64_t)row * nseg + s;
        const uint64_t pos = tile_off[t] + rowpos[rsn] + (k - rs - segoff[rsn]);
        tval[pos] = val[k];
        tcol[pos] = (uint16_t)(c & ((1u << wshift) - 1u));
        trow[pos] = (uint16_t)i;
    }
}

// Rows must be sorted by column for the segment runs to be contiguous.
__global__ __launch_bounds__(kWG) void rows_sorted_kernel(const uint64_t *__restrict__ rowptr,
                                                          const uint32_t *__restrict__ col, uint32_t rows,
                                                          uint32_t *__restrict__ bad) {
    const uint32_t row = blockIdx.x * (kWG / 64) + threadIdx.x / 64;
    const uint32_t lane = threadIdx.x % 64;
    if (row >= rows) return;
    for (uint64_t k = rowptr[row] + 1 + lane; k < rowptr[row + 1]; k += 64)
        if (col[k] < col[k - 1]) atomicOr(bad, 2u);
}

inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }
constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

size_t lds_bytes(uint32_t W, uint32_t R, uint32_t span) { return ((size_t)W + R) * 4 + (size_t)span * 16; }

template <int E, int W>
hipError_t launch_w(const TilePlan &p, const TiledMatrix &tm, const TileWork &wk, const float *x, float *part,
                    hipStream_t s) {
    hipLaunchKernelGGL((blocked_spmv_kernel<E, W>), dim3(wk.nwg), dim3(kWG), lds_bytes(W, p.R, wk.span), s, p.R,
                       wk.span, wk.ntl, wk.wl, wk.tiles, tm.tval, tm.tcol, tm.trow, x, part);
    return hipGetLastError();
}

template <int E>
hipError_t launch_e(const TilePlan &p, const TiledMatrix &tm, const TileWork &wk, const float *x, float *part,
                    hipStream_t s) {
    switch (p.W) {
        case 1024: return launch_w<E, 1024>(p, tm, wk, x, part, s);
        case 2048: return launch_w<E, 2048>(p, tm, wk, x, part, s);
        case 4096: return launch_w<E, 4096>(p, tm, wk, x, part, s);
        default: return launch_w<E, 8192>(p, tm, wk, x, part, s);
    }
}

uint64_t padded_capacity(const TilePlan &p) { return p.nnz + 7 * p.ntiles + 16; }

size_t scan_bytes(const TilePlan &p) {
    size_t scan = 0;
    (void)rocprim::exclusive_scan(nullptr, scan, (const uint64_t *)nullptr, (uint64_t *)nullptr, uint64_t(0),
                                  (size_t)p.ntiles + 1, rocprim::plus<uint64_t>());
    return scan;
}

}  // namespace

TilePlan tile_plan(uint32_t rows, uint32_t cols, uint64_t nnz, uint32_t W, uint32_t R) {
    TilePlan p;
    p.rows = rows;
    p.cols = cols;
    p.nnz = nnz;
    p.W = W;
    p.R = R;
    p.nseg = (cols + W - 1) / W;
    p.nblk = (rows + R - 1) / R;
    p.ntiles = (uint64_t)p.nseg * p.nblk;
    return p;
}

bool valid_tile_width(uint32_t W) { return W == 1024 || W == 2048 || W == 4096 || W == 8192; }
bool valid_tile_rows(uint32_t R) { return R == 512 || R == 1024 || R == 2048 || R == 4096 || R == 8192 || R == 16384; }
bool valid_tile_unit(uint32_t E) { return E == 8 || E == 16; }

size_t tiled_lds_bytes(const TilePlan &p, uint32_t span) { return lds_bytes(p.W, p.R, span); }

size_t tiled_matrix_bytes(const TilePlan &p) {
    const uint64_t cap = padded_capacity(p);
    return align256(8 * (p.ntiles + 1)) + align256(4 * p.ntiles) + align256(4 * cap) + 2 * align256(2 * cap);
}

size_t tiled_part_bytes(const TilePlan &p, uint32_t nwg) { return sizeof(float) * (size_t)nwg * p.R; }

TileWorkList tile_work_list(const TilePlan &p, const uint32_t *nz, const uint64_t *off, uint32_t E,
                            uint32_t target_wgs) {
    const uint64_t UE = (uint64_t)kWG * E;
    TileWorkList L;
    L.E = E;
    uint64_t total = 0;
    for (uint64_t t = 0; t < p.ntiles; ++t) total += (nz[t] + UE - 1) / UE;
    const uint64_t per_run = std::max<uint64_t>(1, (total + std::max(1u, target_wgs) - 1) / std::max(1u, target_wgs));
    L.wl_first.assign((size_t)p.nblk + 1, 0);
    std::vector<uint64_t> units;  // per non-empty tile of the current block
    for (uint32_t b = 0; b < p.nblk; ++b) {
        L.wl_first[b] = L.nwg;
        const uint32_t first = L.ntl;
        uint64_t ub = 0;
        units.clear();
        for (uint32_t sg = 0; sg < p.nseg; ++sg) {
            const uint64_t t = (uint64_t)b * p.nseg + sg;
            if (!nz[t]) continue;
            L.tiles.insert(L.tiles.end(), {sg, nz[t], (uint32_t)off[t], (uint32_t)(off[t] >> 32)});
            ++L.ntl;
            units.push_back((nz[t] + UE - 1) / UE);
            ub += units.back();
        }
        // Runs of nearly equal length over the block's units, none touching
        // more than kTileSpanMax tiles.
        uint64_t runs = (ub + per_run - 1) / per_run, left = ub;
        uint32_t ti = 0;
        uint64_t c = 0;
        while (left) {
            const uint64_t want = (left + runs - 1) / runs;
            const uint32_t t0 = ti;
            const uint64_t c0 = c;
            uint64_t got = 0;
            uint32_t touched = 0;
            while (got < want && touched < kTileSpanMax) {
                ++touched;
                const uint64_t take = std::min(want - got, units[ti] - c);
                got += take;
                c += take;
                if (c == units[ti]) {
                    ++ti;
                    c = 0;
                }
            }
            L.wl.insert(L.wl.end(), {b, first + t0, (uint32_t)c0, (uint32_t)got});
            ++L.nwg;
            L.span = std::max(L.span, touched);
            left -= got;
            runs = runs > 1 ? runs - 1 : 1;
        }
    }
    L.wl_first[p.nblk] = L.nwg;
    return L;
}

size_t tile_build_scratch_bytes(const TilePlan &p) {
    return 2 * align256(4ull * p.rows * p.nseg) + align256(8 * (p.ntiles + 1)) + align256(scan_bytes(p)) + 256;
}

TiledMatrix tiled_matrix_view(const TilePlan &p, void *mem) {
    const uint64_t cap = padded_capacity(p);
    TiledMatrix t;
    char *c = static_cast<char *>(mem);
    t.tile_off = reinterpret_cast<uint64_t *>(c);
    c += align256(8 * (p.ntiles + 1));
    t.tile_nz = reinterpret_cast<uint32_t *>(c);
    c += align256(4 * p.ntiles);
    t.tval = reinterpret_cast<float *>(c);
    c += align256(4 * cap);
    t.tcol = reinterpret_cast<uint16_t *>(c);
    c += align256(2 * cap);
    t.trow = reinterpret_cast<uint16_t *>(c);
    return t;
}

hipError_t launch_tile_build(const TilePlan &p, const uint64_t *rowptr, const uint32_t *col, const float *val,
                             void *mem, void *scratch, uint32_t *flags_host, hipStream_t s) {
    TiledMatrix tm = tiled_matrix_view(p, mem);
    const uint64_t cap = padded_capacity(p);
    char *c = static_cast<char *>(scratch);
    uint32_t *cnt = reinterpret_cast<uint32_t *>(c);
    c += align256(4ull * p.rows * p.nseg);
    uint32_t *rowpos = reinterpret_cast<uint32_t *>(c);
    c += align256(4ull * p.rows * p.nseg);
    uint64_t *tile_padded = reinterpret_cast<uint64_t *>(c);
    c += align256(8 * (p.ntiles + 1));
    size_t sb = scan_bytes(p);
    void *scan_tmp = c;
    c += align256(sb);
    uint32_t *bad = reinterpret_cast<uint32_t *>(c);
    uint32_t wshift = 0;
    while ((1u << wshift) < p.W) ++wshift;

    hipError_t e;
    if ((e = hipMemsetAsync(bad, 0, sizeof(uint32_t), s)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(cnt, 0, 4ull * p.rows * p.nseg, s)) != hipSuccess) return e;
    if (p.rows) {
        hipLaunchKernelGGL(rows_sorted_kernel, dim3(blocks_for(p.rows, kWG / 64)), dim3(kWG), 0, s, rowptr, col,
                           p.rows, bad);
        hipLaunchKernelGGL(tile_count_kernel, dim3(blocks_for(p.rows, kWG / 64)), dim3(kWG), 0, s, rowptr, col,
                           p.rows, p.nseg, wshift, cnt);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // The scatter is only correct (and only in bounds) for sorted rows:
    // check the flag first.
    if ((e = hipMemcpyAsync(flags_host, bad, sizeof(uint32_t), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    if (*flags_host) return hipSuccess;
    hipLaunchKernelGGL(tile_rowpos_kernel, dim3(blocks_for(p.ntiles, kWG)), dim3(kWG), 0, s, cnt, p.rows, p.R,
                       p.nseg, p.ntiles, rowpos, tm.tile_nz, tile_padded);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipMemsetAsync(tile_padded + p.ntiles, 0, sizeof(uint64_t), s)) != hipSuccess) return e;
    if ((e = rocprim::exclusive_scan(scan_tmp, sb, tile_padded, tm.tile_off, uint64_t(0), (size_t)p.ntiles + 1,
                                     rocprim::plus<uint64_t>(), s)) != hipSuccess)
        return e;
    // Padding entries: value 0 (their keys are masked by the entry count).
    if ((e = hipMemsetAsync(tm.tval, 0, 4 * cap, s)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(tm.tcol, 0, 2 * cap, s)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(tm.trow, 0, 2 * cap, s)) != hipSuccess) return e;
    if (p.rows) {
        hipLaunchKernelGGL(row_segscan_kernel, dim3(blocks_for(p.rows, kWG)), dim3(kWG), 0, s, cnt, p.rows, p.nseg);
        hipLaunchKernelGGL(tile_scatter_kernel, dim3(blocks_for(p.rows, kWG / 64)), dim3(kWG), 0, s, rowptr, col, val,
                           p.rows, p.R, p.nseg, wshift, cnt, rowpos, tm.tile_off, tm.tval, tm.tcol, tm.trow);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipStreamSynchronize(s);
}

hipError_t launch_tiled_spmv(const TilePlan &p, const TiledMatrix &tm, const TileWork &wk, const float *x,
                             float *part, Epi epi, const EpiArgs &ea, hipStream_t s) {
    if (p.rows == 0) return hipSuccess;
    if (wk.nwg) {
        hipError_t e = (wk.E == 8) ? launch_e<8>(p, tm, wk, x, part, s) : launch_e<16>(p, tm, wk, x, part, s);
        if (e != hipSuccess) return e;
    }
    uint32_t rshift = 0;
    while ((1u << rshift) < p.R) ++rshift;
    const dim3 grid(blocks_for(p.rows, kWG)), block(kWG);
    switch (epi) {
        case EPI_PLAIN:
            hipLaunchKernelGGL(blocked_reduce_kernel<EPI_PLAIN>, grid, block, 0, s, part, wk.wl_first, rshift, p.rows, ea);
            break;
        case EPI_RATIO:
            hipLaunchKernelGGL(blocked_reduce_kernel<EPI_RATIO>, grid, block, 0, s, part, wk.wl_first, rshift, p.rows, ea);
            break;
        default:
            hipLaunchKernelGGL(blocked_reduce_kernel<EPI_UPDATE>, grid, block, 0, s, part, wk.wl_first, rshift, p.rows,
                               ea);
            break;
    }
    return hipGetLastError();
}

}  // namespace mlem
