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
//   order.  Inside a tile the entries are sorted by (row, column), so every
//   row still sums its entries in the original column order.  Entries come
//   in groups of 8; a group never holds a row step (key difference between
//   neighbours) above 3: a row that starts further away starts a new group
//   and the rest of the old one is padding (value 0, column W, whose LDS
//   slot holds 0).  A tile is cut into units of 2048 entries (the last one
//   shorter, a multiple of 8); unit u of n entries is stored contiguously as
//     n f32 values: entries 0-3 of every group of 8, then entries 4-7 (so
//     each 16-byte load of a wave reads 1 KiB of consecutive bytes),
//     then n/8 group words of 16 bytes:
//       bits  0-15  row - b*R of the group's first entry
//       bits 16-29  2-bit row steps of entries 1..7
//       bits 32-..  8 columns - s*W, 11 bits (W = 1024) or 12 bits (W = 2048)
//   so an entry costs 6 bytes (CSR: 8) and a unit is one 12 KiB stream.
//   tile_off u64[ntiles+1] (entry offsets, padded to 8), tile_nz u32[ntiles]
//   (entries including padding).
//   Quad groups in block streams (the kernel for several right-hand sides,
//   stream_spmv_kernel): every row of a tile starts at a multiple of 4
//   entries, so each half of a group (a quad) holds one row; the group
//   word's bits 0-15 hold the row of entries 0-3, bits 16-31 the row of
//   entries 4-7 (the columns as above).  A row then ends only after entry 3
//   or 7 of a work-item, which is where the accumulator is touched; the
//   price is the padding to whole quads (about 1.5 entries per row
//   segment).  There are no units: the groups of a row block (its tiles in
//   segment order) are stored as one stream, values then group words.
//
// Work: work-item k owns entries [8k, 8k+8) of a unit (two 16-byte value
// loads, one group word).  The units of a row block (its non-empty tiles in
// segment order) are split into runs of nearly equal length, one run per
// workgroup, so every workgroup streams about the same number of bytes
// whatever the tiles' fill (a work list built on the host from the tile
// sizes).  Row ends are where the row key changes.  A row split between
// work-items is finished by the work-item holding its last entry, which adds
// the carries of the work-items before it (back to the previous row end: a
// segmented scan across the wave in registers, then the waves' totals from
// LDS); so each row that has entries in a unit is
// updated by exactly one work-item per unit, and the LDS accumulator of the
// block's rows needs no atomics (a read and a write, ordered by the unit's
// barriers).  The loads of the next unit (and, for a new tile, its x
// segment) are issued before the current unit is computed.  Workgroup w
// writes its partials to part[w][0..R); a reduction kernel adds the
// workgroups of each row block in order and applies the epilogue (ratio or
// update).  Bitwise reproducible.
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cstdlib>
#include <cstring>

#include <hip/hip_fp16.h>

#include "launch.hpp"

namespace mlem {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

template <typename T>
__device__ __forceinline__ T ldnt(const T *p) {
    return __builtin_nontemporal_load(p);
}

constexpr uint32_t kE = kTileE;          // entries per work-item of a unit
constexpr uint32_t kSNT = 512;           // groups per window of stream_spmv_kernel
constexpr uint32_t kQuadAlign = 4;       // quad tiles are padded to a multiple of this many groups
constexpr uint32_t kRunWindows = 64;     // windows of a stream run at most
constexpr uint32_t kUE = kTileUnit;      // entries per unit
constexpr uint32_t kNoKey = 0x10000u;    // past every u16 row key
constexpr uint32_t kMaxRowStep = 3;      // largest key step inside a group of 8 entries
constexpr uint16_t kPadRow = 0xffffu;    // build scratch: position holds padding
static_assert(kE == 8 && kUE == kWG * kE, "one group word per work-item");

template <int W>
constexpr int col_bits() {
    return W == 512 ? 10 : W == 1024 ? 11 : 12;  // columns 0..W (W = padding)
}

// Column j of a group word: bits [32 + j*CB, 32 + (j+1)*CB).  The field
// extract is written as v_bfe_u32 (after a v_alignbit_b32 for a field that
// crosses a dword), so that with the address arithmetic that follows
// (v_lshl_add_u32) a column takes two or three instructions; the compiler
// turns an extract followed by a shift into a shift, a mask and an add,
// three or four.
template <int O, int N>
__device__ __forceinline__ uint32_t bfe_u32(uint32_t w) {
    uint32_t r;
    asm("v_bfe_u32 %0, %1, %2, %3" : "=v"(r) : "v"(w), "n"(O), "n"(N));
    return r;
}
template <int CB, int J>
__device__ __forceinline__ uint32_t group_col_j(const u32x4 &w) {
    constexpr int bit = 32 + J * CB, d = bit / 32, o = bit % 32;
    const uint32_t lo = d == 1 ? w.y : d == 2 ? w.z : w.w;
    if constexpr (o + CB <= 32) {
        return bfe_u32<o, CB>(lo);
    } else {
        const uint32_t hi = d == 1 ? w.z : w.w;
        return bfe_u32<0, CB>(__builtin_amdgcn_alignbit(hi, lo, o));
    }
}
template <int CB>
__device__ __forceinline__ uint32_t group_col(const u32x4 &w, int j) {
    switch (j) {
        case 0: return group_col_j<CB, 0>(w);
        case 1: return group_col_j<CB, 1>(w);
        case 2: return group_col_j<CB, 2>(w);
        case 3: return group_col_j<CB, 3>(w);
        case 4: return group_col_j<CB, 4>(w);
        case 5: return group_col_j<CB, 5>(w);
        case 6: return group_col_j<CB, 6>(w);
        default: return group_col_j<CB, 7>(w);
    }
}

// A run position: staged tile ti (index into the workgroup's tile list),
// unit c of that tile, n entries in the unit.
struct Unit {
    uint32_t ti, c, n;
    bool valid;
};

// Staged tile i: the same for every work-item, so kept in scalar registers
// (unit bookkeeping and load addresses then run on the scalar unit).
__device__ __forceinline__ uint4 uniform_tile(const uint4 *tl, uint32_t i) {
    const uint4 t = tl[i];
    return make_uint4(__builtin_amdgcn_readfirstlane(t.x), __builtin_amdgcn_readfirstlane(t.y),
                      __builtin_amdgcn_readfirstlane(t.z), __builtin_amdgcn_readfirstlane(t.w));
}

__device__ __forceinline__ Unit make_unit(const uint4 *tl, uint32_t ti, uint32_t c, bool valid) {
    Unit u;
    u.ti = ti;
    u.c = c;
    u.valid = valid;
    const uint32_t tj = valid ? ti : 0u;  // ti < span whenever valid
    u.n = valid ? min(kUE, uniform_tile(tl, tj).y - c * kUE) : 0u;
    return u;
}

// Unsigned 16-bit float with EB exponent bits: the f32 with the same
// exponent and mantissa bits (the top binade holds the row maximum, 1.0).
template <int EB>
__device__ __forceinline__ float uf16_lo(uint32_t w) {  // bits 0-15 of w
    constexpr int SH = 7 + EB, EMAX = (1 << EB) - 1;
    // The extract as written (the compiler turns a 16-bit extract at bit 0
    // and the shift into a shift, a mask and an or: three instructions
    // where these are two).
    uint32_t c;
    asm("v_bfe_u32 %0, %1, 0, 16" : "=v"(c) : "v"(w));
    return __uint_as_float((c << SH) + ((uint32_t)(127 - EMAX) << 23));
}
template <int EB>
__device__ __forceinline__ float uf16_hi(uint32_t w) {  // bits 16-31 of w
    constexpr int SH = 7 + EB, EMAX = (1 << EB) - 1;
    return __uint_as_float(((w >> 16) << SH) + ((uint32_t)(127 - EMAX) << 23));
}

// Value j (0..7) of a work-item's group from its loaded 16 bytes: f32
// values come in two halves (entries 0-3, then 4-7), 16-bit values in one
// load, 8 to a work-item.
template <int VT>
__device__ __forceinline__ float group_val(const f32x4 &h0, const f32x4 &h1, const u32x4 &q, int j) {
    if constexpr (VT == VAL_F32) {
        return j < 4 ? h0[j] : h1[j - 4];
    } else if constexpr (VT == VAL_BF16) {
        const uint32_t wv = q[j / 2];
        return __uint_as_float((j & 1) ? (wv & 0xffff0000u) : (wv << 16));
    } else if constexpr (is_uf16(VT)) {
        const uint32_t wv = q[j / 2];
        return (j & 1) ? uf16_hi<kUf16ExpBits>(wv) : uf16_lo<kUf16ExpBits>(wv);
    } else {
        const uint32_t wv = q[j / 2];
        return __half2float(__ushort_as_half((unsigned short)((j & 1) ? (wv >> 16) : (wv & 0xffffu))));
    }
}

template <int W, int VT, int BC>
constexpr int min_waves() {
    return BC == 1 && VT == VAL_F32 ? 7 : 1;
}

// Units whose loads are in flight per workgroup (the kernel takes any depth):
// eight workgroups per CU fit in LDS, one unit each keeps HBM busy.
template <int W, int VT, int BC>
constexpr int pipeline_depth() {
    return 1;
}

// Lane i of the wave gets v from lane i - (shift given by the DPP control
// CTRL) of the rows in ROWM, 0 where there is none.
template <int CTRL, int ROWM>
__device__ __forceinline__ uint32_t dpp_get(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWM, 0xf, false);
}

// One step of a segmented inclusive scan across the wave: (v, f) combined
// with the running value of an earlier lane, where f marks that a segment
// starts at or before this lane (the earlier part then does not count).
// The same for floats, `none` where there is no such lane.
template <int CTRL, int ROWM, bool ZERO>
__device__ __forceinline__ float dpp_getf(float v, float none) {
    return __uint_as_float(
        (uint32_t)__builtin_amdgcn_update_dpp((int)__float_as_uint(none), (int)__float_as_uint(v), CTRL, ROWM, 0xf, ZERO));
}

// v += (v of the lane DPP control DSTR reads) * m, and then m *= (m of
// that lane), in the rows ROWS: one v_fmac_f32_dpp per value (the compiler
// does not fold a DPP read into an FMA).  Lanes without a source lane, and
// rows outside ROWS, are not written (bound_ctrl off): v and m stay.  The
// compiler cannot see the DPP read inside the asm, so the s_nop 1 before
// each gives it the two wait states a DPP read needs after a VALU write of
// the register it reads.  The other DPP hazard, five wait states after a
// VALU write of EXEC (v_cmpx), cannot arise: no code in this file writes
// EXEC from the VALU (tests/test_kernel_asm.py checks the compiled file).
#define MLEM_SEG_STEP_ASM(DSTR, ROWS)                                                                        \
    do {                                                                                                     \
        _Pragma("unroll") for (int b = 0; b < BC; ++b) asm volatile(                                         \
            "s_nop 1\n\tv_fmac_f32_dpp %0, %0, %1 " DSTR " row_mask:" ROWS " bank_mask:0xf"                     \
            : "+v"(v[b])                                                                                     \
            : "v"(m));                                                                                       \
        asm volatile("s_nop 1\n\tv_mul_f32_dpp %0, %0, %0 " DSTR " row_mask:" ROWS " bank_mask:0xf" : "+v"(m)); \
    } while (0)

// One step of a segmented inclusive scan across the wave: v combined with
// the running value of an earlier lane while this lane's segment is still
// open, m = 1.0 (no segment starts at or before the lane in the range seen
// so far) or 0.0.  v + nv * m is exact for m in {0, 1}, and the masks
// combine by a product.
template <int BC, int CTRL>
__device__ __forceinline__ void seg_scan_step(float (&v)[BC], float &m) {
    if constexpr (CTRL == 0x111)
        MLEM_SEG_STEP_ASM("row_shr:1", "0xf");
    else if constexpr (CTRL == 0x112)
        MLEM_SEG_STEP_ASM("row_shr:2", "0xf");
    else if constexpr (CTRL == 0x114)
        MLEM_SEG_STEP_ASM("row_shr:4", "0xf");
    else if constexpr (CTRL == 0x118)
        MLEM_SEG_STEP_ASM("row_shr:8", "0xf");
    else if constexpr (CTRL == 0x142)
        MLEM_SEG_STEP_ASM("row_bcast:15", "0xa");
    else {
        static_assert(CTRL == 0x143, "scan steps: row_shr 1, 2, 4, 8, row_bcast 15, 31");
        MLEM_SEG_STEP_ASM("row_bcast:31", "0xc");
    }
}
#undef MLEM_SEG_STEP_ASM

// One right-hand side.  wl[w] = {row block, first tile (index into tiles),
// first unit in it, units}; tiles[i] = {segment, entries, entry offset lo,
// hi} of the non-empty tiles; part [runs][R].  (The template keeps BC for
// the shared layout arithmetic; it is 1.)
// order[blockIdx.x] = the run this workgroup takes (null: run blockIdx.x);
// the run's index, not the launch position, places its partial sums.
template <int W, int VT, int BC, int D>
__global__ __launch_bounds__(kWG, (min_waves<W, VT, BC>())) void blocked_spmv_kernel(
    uint32_t R, uint32_t span, uint32_t ntl, const uint4 *__restrict__ wl, const uint32_t *__restrict__ order,
    const uint4 *__restrict__ tiles, const uint8_t *__restrict__ tdata, const float *__restrict__ x,
    float *__restrict__ part) {
    constexpr int VB = VT == VAL_F32 ? 4 : 2;  // bytes per stored value
    constexpr int XF = W * BC / kWG;           // floats of an x segment per work-item
    constexpr int CB = col_bits<W>();
    static_assert(BC == 1, "several right-hand sides: stream_spmv_kernel");
    static_assert(XF % 4 == 0, "x segments move in 16-byte pieces");
    extern __shared__ __attribute__((aligned(16))) float smem[];
    constexpr int NW = kWG / 64;        // waves
    __shared__ uint32_t s_k0[kWG + 1];  // first key of every work-item (+ sentinel)
    __shared__ float s_wt[NW];          // carry out of every wave
    __shared__ uint32_t s_wh[NW];       // the wave ends a row
    constexpr uint32_t kDummy = BC == 1 ? kWG : 0;               // dummy rows (one right-hand side only)
    float *xs = smem;                                            // [(W + 4) * BC]; slot W = 0 for padding
    float *acc = xs + (W + 4) * BC;                              // [(R + kDummy) * BC]: rows, then dummies
    uint4 *tl = reinterpret_cast<uint4 *>(acc + (R + kDummy) * BC);  // [span]

    const uint32_t tid = threadIdx.x, lane = tid & 63u, wv = tid >> 6;
    const uint32_t run = order ? order[blockIdx.x] : blockIdx.x;
    const uint4 w = wl[run];
    for (uint32_t i = tid; i < R * BC; i += kWG) acc[i] = 0.0f;
    for (uint32_t i = tid; i < span; i += kWG)
        if (w.y + i < ntl) tl[i] = tiles[w.y + i];
    if (tid < BC) xs[W * BC + tid] = 0.0f;
    if (tid == 0) s_k0[kWG] = kNoKey;
    __syncthreads();

    // Software pipeline: the loads of the next D units are in flight while a
    // unit is computed (D register sets, used in turn).  `issued` counts the
    // units of the run whose loads were issued.
    uint32_t issued = w.w > 0 ? 1u : 0u;
    auto advance = [&](const Unit &u) {  // the unit after u in the run
        if (!u.valid || issued >= w.w) return make_unit(tl, 0, 0, false);
        const bool more_in_tile = (u.c + 1) * kUE < uniform_tile(tl, u.ti).y;
        ++issued;
        return make_unit(tl, more_in_tile ? u.ti : u.ti + 1, more_in_tile ? u.c + 1 : 0u, true);
    };

    f32x4 pv[D][2] = {};
    u32x4 pq[D] = {}, pg[D] = {};
    f32x4 px[D][XF / 4] = {};
    // Loads of unit u into register set k; an invalid u re-reads the start of
    // the run's first tile and lanes past the unit's end re-read lane 0's
    // entries (every address stays inside a unit, so the wait counts are
    // exact).
    // x only for a unit that is staged: the run's first (need_x) or one that
    // starts a tile (a tile spans about three units on the 3D scanner matrix,
    // and loading x for every unit was a third of the workgroup's loads).
    auto prefetch = [&](int k, const Unit &u, bool need_x) {
        const uint4 t = uniform_tile(tl, u.valid ? u.ti : 0u);
        const uint64_t off = ((uint64_t)t.w << 32) | t.z;
        const uint32_t c = u.valid ? u.c : 0u;
        const uint32_t n8 = (min(kUE, t.y - c * kUE) + 7u) & ~7u;
        const uint8_t *base = tdata + (VB + 2) * (off + (uint64_t)c * kUE);
        const uint32_t le = (tid * kE < (u.valid ? u.n : 1u)) ? tid * kE : 0u;
        if constexpr (VT == VAL_F32) {
            pv[k][0] = ldnt(reinterpret_cast<const f32x4 *>(base) + le / 8);
            pv[k][1] = ldnt(reinterpret_cast<const f32x4 *>(base + 2 * n8) + le / 8);
        } else {
            pq[k] = ldnt(reinterpret_cast<const u32x4 *>(base) + le / 8);
        }
        pg[k] = ldnt(reinterpret_cast<const u32x4 *>(base + VB * n8) + le / 8);
        // x is padded to whole segments (xlen >= nseg * W): no clamping.
        if (need_x || (u.valid && u.c == 0)) {
            const float *xb = x + (uint64_t)t.x * W * BC;
#pragma unroll
            for (int q = 0; q < XF / 4; ++q) px[k][q] = *(reinterpret_cast<const f32x4 *>(xb) + tid + kWG * q);
        }
    };

    Unit su[D];  // the unit whose loads are in register set k
    Unit iss = make_unit(tl, 0, w.z, w.w > 0);  // the last unit issued
#pragma unroll
    for (int k = 0; k < D; ++k) {
        if (k) iss = advance(iss);
        su[k] = iss;
        prefetch(k, iss, k == 0);
    }
    bool fresh = true;  // first unit of the run: its x segment is not staged yet
    // Every round runs all D stages, so the loop has one exit and the loads'
    // wait counts stay exact; the stages past the run's end (at most D-1, in
    // the last round) see units without entries and change nothing.
    const uint32_t rounds = (w.w + D - 1) / D;
    for (uint32_t rd = 0; rd < rounds; ++rd) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const Unit cur = su[k];
            const f32x4 cv0 = pv[k][0], cv1 = pv[k][1];
            const u32x4 cq = pq[k], cg = pg[k];
            f32x4 cx[XF / 4];
#pragma unroll
            for (int q = 0; q < XF / 4; ++q) cx[q] = px[k][q];
            // The next unit's loads go into set k.  One set: issued now, in
            // flight while this unit is computed.  Several: issued once this
            // unit's registers are used up (after the walk), in flight while
            // the next D-1 units are computed (loading into registers that
            // still hold live data would make the compiler copy them, and a
            // copy waits for the loads).
            auto issue_next = [&]() {
                iss = advance(iss);
                su[k] = iss;
                prefetch(k, iss, false);
            };
            if constexpr (D == 1) issue_next();

            const bool lane_valid = tid * kE < cur.n;  // entries [n, n8) are padding
            // First row key of every work-item, for the row ends at its last entry.
            const uint32_t key0 = lane_valid ? (cg.x & 0xffffu) : R + tid;
            s_k0[tid] = lane_valid ? key0 : kNoKey;
            if (cur.c == 0 || fresh) {  // a new tile: its x segment
#pragma unroll
                for (int q = 0; q < XF / 4; ++q) *(reinterpret_cast<f32x4 *>(xs) + tid + kWG * q) = cx[q];
            }
            fresh = false;
            __syncthreads();
            if constexpr (BC == 1) {
                float a[1];  // the part of the last row that continues in the next work-item
                // Row steps (the builder checks that every key stays below R and
                // every column at most W).  A work-item past the unit's end reads
                // lane 0's entries: it has the key R + tid (its own dummy
                // accumulator row) and no row steps, so whatever it adds lands
                // there.
                const uint32_t word = lane_valid ? cg.x : 0u;
                uint32_t step[kE];  // step[j]: key of entry j minus key of entry j-1
#pragma unroll
                for (int j = 1; j < 8; ++j) step[j] = (word >> (14 + 2 * j)) & 3u;
                uint32_t key7 = key0;
#pragma unroll
                for (int j = 1; j < 8; ++j) key7 += step[j];
                // Row ends: entry j < 7 ends its row where step[j+1] != 0, entry 7
                // where the next work-item starts another row.  Every row that
                // ends here gets its partial added to the accumulator (this
                // work-item is the only writer of those rows in this unit); the
                // other entries read and write the dummy row R + tid, so there
                // are no branches.  All reads come before any write.
                const bool end7 = key7 != s_k0[tid + 1];
                uint32_t slot[kE];
                {
                    uint32_t k = key0;
#pragma unroll
                    for (int m = 0; m < (int)kE; ++m) {
                        const bool e = m < 7 ? step[m + 1] != 0u : end7;
                        slot[m] = e ? k : R + tid;
                        if (m < 7) k += step[m + 1];
                    }
                }
                float old[kE];
#pragma unroll
                for (int m = 0; m < (int)kE; ++m) old[m] = acc[slot[m]];
                // Products (padding reads column W, whose slot holds 0), the
                // running row sum (restarted after each row end) added into the
                // slots.
                a[0] = 0.0f;
#pragma unroll
                for (int m = 0; m < (int)kE; ++m) {
                    const float v = group_val<VT>(cv0, cv1, cq, m);
                    const bool end = m < 7 ? step[m + 1] != 0u : end7;
                    a[0] += v * xs[group_col<CB>(cg, m)];
                    old[m] += a[0];
                    a[0] = end ? 0.0f : a[0];
                }
#pragma unroll
                for (int m = 0; m < (int)kE; ++m) acc[slot[m]] = old[m];
                uint32_t ends = end7 ? 1u : 0u;
#pragma unroll
                for (int j = 1; j < 8; ++j) ends |= step[j];
                if constexpr (D > 1) issue_next();
                // a[0] is now the part of the last row that continues in the
                // next work-item.  The row of entry 0 may have started in
                // earlier work-items: when it ends here, it gets their carries
                // (back to the last one with a row end, whose carry belongs to
                // this row too): a segmented scan over the carries shifted by
                // one work-item, across the wave in registers (stopping once
                // no lane that uses its result needs more), then across the
                // waves from their totals in LDS.
                float sc[1] = {dpp_getf<0x138, 0xf, true>(a[0], 0.0f)};  // wave_shr:1
                float mo = dpp_getf<0x138, 0xf, false>(ends ? 0.0f : 1.0f, 1.0f);
                seg_scan_step<1, 0x111>(sc, mo);  // row_shr:1
                seg_scan_step<1, 0x112>(sc, mo);  // row_shr:2
                seg_scan_step<1, 0x114>(sc, mo);  // row_shr:4
                const bool uses = ends || lane == 63u;
                if (__builtin_amdgcn_ballot_w64(uses && mo != 0.0f && lane >= 8u)) {
                    seg_scan_step<1, 0x118>(sc, mo);  // row_shr:8
                    if (__builtin_amdgcn_ballot_w64(uses && mo != 0.0f && lane >= 16u)) {
                        seg_scan_step<1, 0x142>(sc, mo);  // row_bcast:15
                        seg_scan_step<1, 0x143>(sc, mo);  // row_bcast:31
                    }
                }
                if (lane == 63u) {
                    s_wt[wv] = ends ? a[0] : sc[0] + a[0];
                    s_wh[wv] = (mo == 0.0f || ends) ? 1u : 0u;
                }
                __syncthreads();
                if (ends) {
                    float cin = sc[0];
                    if (mo != 0.0f) {  // the row started before this wave: nearest wave first
                        for (int u = (int)wv - 1; u >= 0; --u) {
                            cin += s_wt[u];
                            if (s_wh[u]) break;
                        }
                    }
                    acc[key0] += cin;
                }
            }
        }
    }
    __syncthreads();
    float *pw = part + (uint64_t)run * R * BC;
    for (uint32_t i = tid; i < R * BC; i += kWG) pw[i] = acc[i];
}

// Several right-hand sides: quad groups in block-stream order (the groups of
// a row block's tiles one after another, values then group words; every
// tile padded to a multiple of kQuadAlign groups).  A window is kSNT
// consecutive groups of the stream from at most two tiles (A, then B from
// group `split` on); work-item t of an NT-wide workgroup takes the GPW =
// kSNT / NT consecutive groups [t*GPW, t*GPW + GPW) of it, all in one tile
// (window starts, ends and splits are multiples of kQuadAlign groups).  The
// x segments of the tiles sit in two LDS slots by the parity of the tile's
// place in the block's tile list, so a window does not stop at a tile
// boundary.
// wl[w] = {row block, first window, windows}; win[i] = {first group,
// groups | split << 16 | slot of A << 31, segment of A, segment of B}.


template <int W, int VT, int BC, int NT>
__global__ __launch_bounds__(NT) void stream_spmv_kernel(uint32_t R, uint32_t nseg,
                                                         const uint64_t *__restrict__ tile_off,
                                                         const uint4 *__restrict__ wl,
                                                         const uint32_t *__restrict__ order,
                                                         const uint4 *__restrict__ win,
                                                         const uint8_t *__restrict__ tdata,
                                                         const float *__restrict__ x, float *__restrict__ part) {
    constexpr int GPW = kSNT / NT;             // groups per work-item and window
    constexpr int NQ = 2 * GPW;                // quads per work-item
    constexpr int VB = VT == VAL_F32 ? 4 : 2;  // bytes per stored value
    constexpr int XF = W * BC / NT;            // floats of an x segment per work-item
    constexpr int CB = col_bits<W>();
    constexpr uint32_t XS = (W + 4) * BC;      // floats per x slot (column W holds 0)
    constexpr int NW = NT / 64;                // waves
    static_assert(kSNT % NT == 0 && kQuadAlign % GPW == 0, "a work-item's groups stay in one tile");
    static_assert(BC % 4 == 0 && XF % 4 == 0, "x segments move in 16-byte pieces");
    extern __shared__ __attribute__((aligned(16))) float smem[];
    __shared__ uint32_t s_k0[NT + 1];                           // first key of every work-item
    __shared__ __attribute__((aligned(16))) float s_wt[NW][BC];  // carry out of every wave
    __shared__ uint32_t s_wh[NW];                               // the wave ends a row
    __shared__ uint4 s_win[kRunWindows];                        // the run's windows
    float *xs = smem;                                           // [2][XS]
    float *acc = xs + 2 * XS;                                   // [R * BC]

    const uint32_t tid = threadIdx.x, lane = tid & 63u, wv = tid >> 6;
    const uint32_t g0 = tid * GPW;  // the work-item's first group in a window
    const uint32_t run = order[blockIdx.x];  // workgroups start in column order, runs keep their place
    const uint4 w = wl[run];
    const uint64_t e0 = tile_off[(uint64_t)w.x * nseg], e1 = tile_off[(uint64_t)(w.x + 1) * nseg];
    const uint64_t G = (e1 - e0) / 8;  // groups of the row block
    const uint8_t *vbase = tdata + (uint64_t)(VB + 2) * e0;
    const uint8_t *gbase = vbase + (uint64_t)VB * 8 * G;
    for (uint32_t i = tid; i < R * BC; i += NT) acc[i] = 0.0f;
    for (uint32_t i = tid; i < w.z; i += NT) s_win[i] = win[w.y + i];
    if (tid < BC) {
        xs[W * BC + tid] = 0.0f;
        xs[XS + W * BC + tid] = 0.0f;
    }
    if (tid == 0) s_k0[NT] = kNoKey;

    uint32_t staged[2] = {~0u, ~0u};  // segment held by each x slot
    // Loads of window c (its groups; the x segments of its tiles that are
    // not staged yet: needA / needB).  Work-items past the window's end
    // re-read its first group.
    f32x4 pv0[GPW], pv1[GPW];
    u32x4 pq[GPW], pg[GPW];
    f32x4 pxa[XF / 4], pxb[XF / 4];
    auto load = [&](const uint4 &c, bool needA, bool needB) {
        const uint32_t ng = c.y & 0x3ffu;
        const bool valid = g0 < ng;
        const uint64_t gb = c.x + (valid ? g0 : 0u);
#pragma unroll
        for (int j = 0; j < GPW; ++j) {
            const uint64_t g = gb + (valid ? j : 0);
            if constexpr (VT == VAL_F32) {
                pv0[j] = ldnt(reinterpret_cast<const f32x4 *>(vbase) + g);
                pv1[j] = ldnt(reinterpret_cast<const f32x4 *>(vbase + 16 * G) + g);
            } else {
                pq[j] = ldnt(reinterpret_cast<const u32x4 *>(vbase) + g);
            }
            pg[j] = ldnt(reinterpret_cast<const u32x4 *>(gbase) + g);
        }
        // x is padded to whole segments (xlen >= nseg * W): no clamping.
        if (needA) {
            const f32x4 *xa = reinterpret_cast<const f32x4 *>(x + (uint64_t)c.z * W * BC);
#pragma unroll
            for (int q = 0; q < XF / 4; ++q) pxa[q] = xa[tid + NT * q];
        }
        if (needB) {
            const f32x4 *xb = reinterpret_cast<const f32x4 *>(x + (uint64_t)c.w * W * BC);
#pragma unroll
            for (int q = 0; q < XF / 4; ++q) pxb[q] = xb[tid + NT * q];
        }
    };
    auto uniform4 = [](const uint4 &t) {
        return make_uint4(__builtin_amdgcn_readfirstlane(t.x), __builtin_amdgcn_readfirstlane(t.y),
                          __builtin_amdgcn_readfirstlane(t.z), __builtin_amdgcn_readfirstlane(t.w));
    };
    uint4 cw = uniform4(win[w.y]);
    bool needA = true, needB = ((cw.y >> 16) & 0x3ffu) < (cw.y & 0x3ffu);
    load(cw, needA, needB);
    __syncthreads();

    for (uint32_t i = 0; i < w.z; ++i) {
        const uint32_t ng = cw.y & 0x3ffu, split = (cw.y >> 16) & 0x3ffu, slotA = cw.y >> 31;
        f32x4 cv0[GPW], cv1[GPW];
        u32x4 cq[GPW], cg[GPW];
#pragma unroll
        for (int j = 0; j < GPW; ++j) {
            cv0[j] = pv0[j];
            cv1[j] = pv1[j];
            cq[j] = pq[j];
            cg[j] = pg[j];
        }
        const uint4 nw = uniform4(s_win[i + 1 < w.z ? i + 1 : i]);
        // Stage the new x segments (the slots' previous tiles are done: the
        // last barrier of the previous window lies between).
        if (needA) {
#pragma unroll
            for (int q = 0; q < XF / 4; ++q) *(reinterpret_cast<f32x4 *>(xs + slotA * XS) + tid + NT * q) = pxa[q];
            staged[slotA] = cw.z;
        }
        if (needB) {
#pragma unroll
            for (int q = 0; q < XF / 4; ++q)
                *(reinterpret_cast<f32x4 *>(xs + (slotA ^ 1u) * XS) + tid + NT * q) = pxb[q];
            staged[slotA ^ 1u] = cw.w;
        }
        const bool lane_valid = g0 < ng;
        const bool inA = g0 < split;
        // Quad 2j holds entries 0-3 of group j (row key bits 0-15), quad
        // 2j+1 entries 4-7 (bits 16-31); keys do not decrease inside a tile.
        uint32_t kq[NQ];
#pragma unroll
        for (int j = 0; j < GPW; ++j) {
            kq[2 * j] = cg[j].x & 0xffffu;
            kq[2 * j + 1] = cg[j].x >> 16;
        }
        s_k0[tid] = lane_valid ? kq[0] : kNoKey;
        // The next window's loads, in flight while this one is computed.
        if (i + 1 < w.z) {
            const uint32_t sa = nw.y >> 31;
            needA = staged[sa] != nw.z;
            needB = ((nw.y >> 16) & 0x3ffu) < (nw.y & 0x3ffu) && staged[sa ^ 1u] != nw.w;
            load(nw, needA, needB);
            cw = nw;
        }
        __syncthreads();
        const float *xw = xs + (inA ? slotA : slotA ^ 1u) * XS;
        // Every quad is summed on its own (padding reads column W, whose
        // slot holds 0); work-items past the window's end skip the sums.
        // Work-items past the window's end sum their re-read first group
        // too (finite values, columns within the slot) and drop the sums:
        // they end no row and carry 0.
        // A quad's four x reads are issued together, then its products.
        float sq[NQ][BC];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            f32x4 t4[4][BC / 4];
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const float *xg = xw + group_col<CB>(cg[q / 2], 4 * (q & 1) + jj) * BC;
#pragma unroll
                for (int qq = 0; qq < BC / 4; ++qq) t4[jj][qq] = *reinterpret_cast<const f32x4 *>(xg + 4 * qq);
            }
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const float v = group_val<VT>(cv0[q / 2], cv1[q / 2], cq[q / 2], 4 * (q & 1) + jj);
#pragma unroll
                for (int qq = 0; qq < BC / 4; ++qq)
#pragma unroll
                    for (int bb = 0; bb < 4; ++bb)
                        sq[q][4 * qq + bb] =
                            jj == 0 ? v * t4[jj][qq][bb] : fmaf(v, t4[jj][qq][bb], sq[q][4 * qq + bb]);
            }
        }
        // A row ends after quad q where the next quad (the next work-item's
        // first, after the last) holds another row.  sq[q] becomes the sum
        // of the quads since the work-item's previous row end.
        const uint32_t knext = s_k0[tid + 1];
        bool endq[NQ];
        bool ends = false;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            endq[q] = lane_valid && kq[q] != (q + 1 < NQ ? kq[q + 1] : knext);
            ends = ends || endq[q];
        }
#pragma unroll
        for (int q = 1; q < NQ; ++q)
#pragma unroll
            for (int c = 0; c < BC; ++c) sq[q][c] = endq[q - 1] ? sq[q][c] : sq[q - 1][c] + sq[q][c];
        // Carry out: the part of the last row that continues in the next
        // work-item.  The carry in of the work-item's first row is the sum
        // of the carries of the work-items since the last one that ends a
        // row (that one's included), or since the window's start: a
        // segmented scan over the carries shifted by one work-item, in
        // registers across the wave, then across the waves from their
        // totals in LDS.
        float a[BC], s[BC];
        const bool open = lane_valid && !endq[NQ - 1];
#pragma unroll
        for (int b = 0; b < BC; ++b) a[b] = open ? sq[NQ - 1][b] : 0.0f;
#pragma unroll
        for (int b = 0; b < BC; ++b) s[b] = dpp_getf<0x138, 0xf, true>(a[b], 0.0f);  // wave_shr:1
        float m = dpp_getf<0x138, 0xf, false>(ends ? 0.0f : 1.0f, 1.0f);
        // The scan stops as soon as no lane that uses its carry in (one that
        // ends a row, and lane 63 for the wave's carry out) still lacks the
        // start of its segment: after the steps of shift 1, 2, 4 a lane has
        // the 7 lanes before it (within its row of 16), after shift 8 its
        // row, after the broadcasts the wave.  Rows of a few quads (the
        // transpose's voxel rows) rarely need more than the first three
        // steps; the later steps would change no lane that is used.
        seg_scan_step<BC, 0x111>(s, m);  // row_shr:1
        seg_scan_step<BC, 0x112>(s, m);  // row_shr:2
        seg_scan_step<BC, 0x114>(s, m);  // row_shr:4
        const bool uses = ends || lane == 63u;
        if (__builtin_amdgcn_ballot_w64(uses && m != 0.0f && lane >= 8u)) {
            seg_scan_step<BC, 0x118>(s, m);  // row_shr:8
            if (__builtin_amdgcn_ballot_w64(uses && m != 0.0f && lane >= 16u)) {
                seg_scan_step<BC, 0x142>(s, m);  // row_bcast:15
                seg_scan_step<BC, 0x143>(s, m);  // row_bcast:31
            }
        }
        const bool f = m == 0.0f;  // a row ends in an earlier lane of this wave
        if (lane == 63) {  // the wave's carry out and whether it ends a row
#pragma unroll
            for (int b = 0; b < BC; ++b) s_wt[wv][b] = ends ? a[b] : s[b] + a[b];
            s_wh[wv] = (f || ends) ? 1u : 0u;
        }
        __syncthreads();
        // Every row that ends here is added to the accumulator once, the
        // work-item's first with the carry in.  A row can end twice in a
        // window that spans two tiles (once in each): the work-items of
        // tile A update the accumulator first, a barrier, then those of
        // tile B.
        auto flush = [&]() {
            float c[BC];
#pragma unroll
            for (int b = 0; b < BC; ++b) c[b] = s[b];
            if (!f) {  // the row started before this wave: the earlier waves' carries, nearest first,
                       // back to the nearest wave that ends a row (its carry out included)
                for (int u = (int)wv - 1; u >= 0; --u) {
#pragma unroll
                    for (int b = 0; b < BC; ++b) c[b] += s_wt[u][b];
                    if (s_wh[u]) break;
                }
            }
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                if (endq[q]) {
                    const uint32_t r = kq[q] * BC;
#pragma unroll
                    for (int qq = 0; qq < BC / 4; ++qq) {
                        f32x4 o = *reinterpret_cast<const f32x4 *>(acc + r + 4 * qq);
#pragma unroll
                        for (int bb = 0; bb < 4; ++bb) o[bb] += c[4 * qq + bb] + sq[q][4 * qq + bb];
                        *reinterpret_cast<f32x4 *>(acc + r + 4 * qq) = o;
                    }
#pragma unroll
                    for (int b = 0; b < BC; ++b) c[b] = 0.0f;
                }
            }
        };
        if (ends && inA) flush();
        if (split < ng) {
            __syncthreads();
            if (ends && !inA) flush();
        }
    }
    __syncthreads();
    float *pw = part + (uint64_t)run * R * BC;
    for (uint32_t i = tid; i < R * BC; i += NT) pw[i] = acc[i];
}

// The epilogue of row r from its sum y over the block's workgroups: the
// row's value scale (16-bit values), then the ratio, the update or the plain
// sum.  Vectors are [rows][BC].  Its inputs are read by reduce_load before
// the partials are summed, so that their memory round trip overlaps the
// sum's instead of following it (a small matrix's reduction is a chain of
// dependent round trips).
template <int BC>
struct ReduceIn {
    float sc;     // the row's value scale (1 without scales)
    float a[BC];  // RATIO: the counts; UPDATE: the image values
    float n;      // UPDATE: the sensitivity
    bool act[BC]; // UPDATE: the right-hand side is active
};

template <int EPI, int BC>
__device__ __forceinline__ ReduceIn<BC> reduce_load(uint32_t r, const float *__restrict__ scale, const EpiArgs &ea) {
    ReduceIn<BC> in;
    in.sc = scale ? scale[r] : 1.0f;
    in.n = 0.0f;
#pragma unroll
    for (int q = 0; q < BC; ++q) {
        const uint64_t k = (uint64_t)r * BC + q;
        in.a[q] = 0.0f;
        in.act[q] = true;
        if constexpr (EPI == EPI_RATIO) in.a[q] = (float)ea.g[k];
        if constexpr (EPI == EPI_UPDATE) {
            in.a[q] = ea.f[k];
            in.act[q] = !ea.active || ea.active[q];
        }
    }
    if constexpr (EPI == EPI_UPDATE) in.n = ea.norm[r];
    return in;
}

template <int EPI, int BC>
__device__ __forceinline__ void reduce_epilogue(uint32_t r, const float (&y)[BC], const ReduceIn<BC> &in,
                                                bool scaled, const EpiArgs &ea) {
#pragma unroll
    for (int q = 0; q < BC; ++q) {
        const uint64_t k = (uint64_t)r * BC + q;
        const float yq = scaled ? y[q] * in.sc : y[q];
        if constexpr (EPI == EPI_PLAIN) {
            ea.out[k] = yq;
        } else if constexpr (EPI == EPI_RATIO) {
            // 16-bit values: the back projection's tiles hold the same codes
            // without the row scales, so the ratio g / (scale * sum) goes on
            // multiplied by the scale, that is g / sum.
            ea.out[k] = (y[q] != 0.0f) ? in.a[q] / y[q] : 0.0f;
        } else {
            const float v = in.a[q] * yq;
            if (in.act[q]) ea.f[k] = (in.n == 0.0f) ? v : v / in.n;
        }
    }
}

// y += the partials of row i (of R = 2^rshift) in workgroups w0, w0 + S,
// ... below we, in that order.  The partials of U workgroups are loaded
// together, then added in order: one memory latency per U workgroups
// instead of one each.
template <int BC, uint32_t S>
__device__ __forceinline__ void sum_partials(float (&y)[BC], const float *__restrict__ part, uint32_t w0,
                                             uint32_t we, uint32_t rshift, uint32_t i) {
    constexpr uint32_t U = BC == 1 ? 16 : (BC == 4 ? 8 : 4);
    uint32_t w = w0;
    for (; w + (U - 1) * S < we; w += U * S) {
        float t[U][BC];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const float *pr = part + ((((uint64_t)(w + u * S)) << rshift) + i) * BC;
#pragma unroll
            for (int q = 0; q < BC; ++q) t[u][q] = pr[q];
        }
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
#pragma unroll
            for (int q = 0; q < BC; ++q) y[q] += t[u][q];
    }
    for (; w < we; w += S) {
        const float *pr = part + (((uint64_t)w << rshift) + i) * BC;
#pragma unroll
        for (int q = 0; q < BC; ++q) y[q] += pr[q];
    }
}

// Row r of block b = r / R sums the partials of the block's workgroups
// wl_first[b] .. wl_first[b+1]-1 in order (for each of the BC right-hand
// sides) and applies the epilogue.
template <int EPI, int BC>
__global__ void blocked_reduce_kernel(const float *__restrict__ part, const uint32_t *__restrict__ wl_first,
                                      uint32_t rshift, uint32_t rows, const float *__restrict__ scale, EpiArgs ea) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const uint32_t b = r >> rshift, i = r & ((1u << rshift) - 1u);
    const ReduceIn<BC> in = reduce_load<EPI, BC>(r, scale, ea);
    float y[BC];
#pragma unroll
    for (int q = 0; q < BC; ++q) y[q] = 0.0f;
    sum_partials<BC, 1>(y, part, wl_first[b], wl_first[b + 1], rshift, i);
    reduce_epilogue<EPI, BC>(r, y, in, scale != nullptr, ea);
}

// The same for blocks with many workgroups (a small matrix spread over the
// whole chip has hundreds per block): wave j of a (64, kSlices) workgroup
// sums workgroups j, j + kSlices, ... of each of 64 consecutive rows, then
// wave 0 adds the kSlices sums in slice order.
constexpr uint32_t kSlices = 8;
constexpr uint32_t kSlicedRuns = 64;  // runs of a row block from which the sliced sum is used
template <int EPI, int BC>
__global__ __launch_bounds__(64 * kSlices) void blocked_reduce_sliced_kernel(
    const float *__restrict__ part, const uint32_t *__restrict__ wl_first, uint32_t rshift, uint32_t rows,
    const float *__restrict__ scale, EpiArgs ea) {
    __shared__ float s_y[kSlices][64][BC];
    const uint32_t x = threadIdx.x, j = threadIdx.y;
    const uint32_t r = blockIdx.x * 64 + x;
    float y[BC];
#pragma unroll
    for (int q = 0; q < BC; ++q) y[q] = 0.0f;
    ReduceIn<BC> in{};
    if (r < rows) {
        if (j == 0) in = reduce_load<EPI, BC>(r, scale, ea);  // wave 0 applies the epilogue
        const uint32_t b = r >> rshift, i = r & ((1u << rshift) - 1u);
        sum_partials<BC, kSlices>(y, part, wl_first[b] + j, wl_first[b + 1], rshift, i);
    }
#pragma unroll
    for (int q = 0; q < BC; ++q) s_y[j][x][q] = y[q];
    __syncthreads();
    if (j != 0 || r >= rows) return;
#pragma unroll
    for (uint32_t jj = 1; jj < kSlices; ++jj)
#pragma unroll
        for (int q = 0; q < BC; ++q) y[q] += s_y[jj][x][q];
    reduce_epilogue<EPI, BC>(r, y, in, scale != nullptr, ea);
}

// ----------------------------------------------------------------- builder
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

// Per tile (one work-item per tile): rowpos[r*nseg+s] = position of row r's
// run in tile (b, s); tile_nz = entries including step padding, padded to 8
// in tile_padded.  Row steps: a row whose key exceeds the previous non-empty
// row's by more than kMaxRowStep starts a new group of 8 (the rest of the
// group is padding), so every key inside a group is its predecessor plus
// 0..3.  Quads: every non-empty row starts at a multiple of 4, and the tile
// is padded to a multiple of kQuadAlign groups.
__global__ void tile_rowpos_kernel(const uint32_t *__restrict__ cnt, uint32_t rows, uint32_t R, uint32_t nseg,
                                   uint64_t ntiles, bool quad, uint32_t *__restrict__ rowpos,
                                   uint32_t *__restrict__ tile_nz, uint64_t *__restrict__ tile_padded) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const uint32_t b = (uint32_t)(t / nseg), s = (uint32_t)(t % nseg);
    uint32_t run = 0, last = 0;
    bool any = false;
    for (uint32_t i = 0; i < R; ++i) {
        const uint64_t r = (uint64_t)b * R + i;
        if (r >= rows) break;
        const uint32_t n = cnt[r * nseg + s];
        if (n) {
            if (quad)
                run = (run + 3u) & ~3u;
            else if (any && i - last > kMaxRowStep)
                run = (run + 7u) & ~7u;
            any = true;
            last = i;
        }
        rowpos[r * nseg + s] = run;
        run += n;
    }
    tile_nz[t] = run;
    const uint32_t align = quad ? 8 * kQuadAlign : 8;
    tile_padded[t] = (run + align - 1) / align * align;
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

// scale[r] = mult x the largest |value| of row r (mult for a row without
// nonzero values): 16-bit tiles store value / scale.  FP16 takes mult =
// 2^-14, so a row's largest value is stored as 16384 and even values 1e-8 of
// it stay normal numbers (11-bit significand); BF16 has the range anyway.
__global__ __launch_bounds__(kWG) void row_absmax_kernel(const uint64_t *__restrict__ rowptr,
                                                         const float *__restrict__ val, uint32_t rows, float mult,
                                                         float *__restrict__ scale) {
    const uint32_t row = blockIdx.x * (kWG / 64) + threadIdx.x / 64;
    const uint32_t lane = threadIdx.x % 64;
    if (row >= rows) return;
    float m = 0.0f;
    for (uint64_t k = rowptr[row] + lane; k < rowptr[row + 1]; k += 64) m = fmaxf(m, fabsf(val[k]));
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if (lane == 0) scale[row] = (m > 0.0f ? m : 1.0f) * mult;
}

// f32 -> bf16 bits, round to nearest even (finite inputs).
__device__ __forceinline__ uint16_t to_bf16(float f) {
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float from_bf16(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// t in [0, 1] -> unsigned 16-bit float with EB exponent bits, mantissa
// rounded to nearest even (every value from 2^-(2^EB - 1) up within 2^-(17 -
// EB) of itself); below the smallest binade: code 0.
__device__ __forceinline__ uint16_t to_uf16(float t, int EB) {
    const int SH = 7 + EB, EMAX = (1 << EB) - 1;
    const uint32_t u = __float_as_uint(t);
    const uint32_t ur = u + ((1u << (SH - 1)) - 1u) + ((u >> SH) & 1u);
    const int32_t c = (int32_t)(ur >> SH) - ((127 - EMAX) << (16 - EB));
    return (uint16_t)(c < 0 ? 0 : c);
}

__device__ __forceinline__ float from_uf16(uint16_t c, int EB) {
    const int SH = 7 + EB, EMAX = (1 << EB) - 1;
    return __uint_as_float(((uint32_t)c << SH) + ((uint32_t)(127 - EMAX) << 23));
}

// 16-bit values, each value / scale of its row rounded to nearest (even).
// Rounding every value on its own keeps each entry within half a unit in the
// last place; carrying the rounding error along the row (error feedback)
// keeps row sums closer but can move a small entry by many times itself,
// which the back projection and the sensitivity of the voxels behind it do
// not forgive (measured: rel-L2 up to 0.13 after 100 iterations on the 3D
// scanner matrix, against at most 3e-3 for FP16 without it).
// Unsigned 16-bit floats take nonnegative values only: bit 8 of *bad for a
// negative one.
__global__ void quantize_rows_kernel(const uint64_t *__restrict__ rowptr, const float *__restrict__ val,
                                     uint32_t rows, const float *__restrict__ scale, uint32_t vt,
                                     uint16_t *__restrict__ q, uint32_t *__restrict__ bad) {
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    const float inv = 1.0f / scale[row];
    bool neg = false;
    for (uint64_t k = rowptr[row]; k < rowptr[row + 1]; ++k) {
        const float t = val[k] * inv;
        if (is_uf16(vt)) {
            neg = neg || t < 0.0f;
            q[k] = to_uf16(fminf(fmaxf(t, 0.0f), 1.0f), kUf16ExpBits);
        } else {
            q[k] = vt == VAL_BF16 ? to_bf16(t) : __half_as_ushort(__float2half_rn(t));
        }
    }
    if (neg && bad) atomicOr(bad, 8u);
}

// Places every entry: its value straight into the unit it falls in, its
// local column and row key into scratch (by entry offset) for the packing.
__global__ __launch_bounds__(kWG) void tile_scatter_kernel(const uint64_t *__restrict__ rowptr,
                                                           const uint32_t *__restrict__ col,
                                                           const float *__restrict__ val, uint32_t rows, uint32_t R,
                                                           uint32_t nseg, uint32_t wshift,
                                                           const uint32_t *__restrict__ segoff,
                                                           const uint32_t *__restrict__ rowpos,
                                                           const uint64_t *__restrict__ tile_off, uint32_t vt,
                                                           bool stream, const uint16_t *__restrict__ qval,
                                                           uint8_t *__restrict__ tdata, uint16_t *__restrict__ ctmp,
                                                           uint16_t *__restrict__ rtmp) {
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
        const uint64_t off = tile_off[t];
        const uint64_t p = rowpos[rsn] + (k - rs - segoff[rsn]);  // position in the tile
        // Entry q of a stretch of n8u entries stored from entry offset `base`:
        // a unit of a tile, or (stream) the whole row block.
        uint64_t base, q, n8u;
        if (stream) {
            base = tile_off[(uint64_t)b * nseg];
            q = off - base + p;
            n8u = tile_off[(uint64_t)(b + 1) * nseg] - base;
        } else {
            const uint64_t u = p / kUE;
            base = off + u * kUE;
            q = p % kUE;
            n8u = std::min<uint64_t>(kUE, tile_off[t + 1] - base);
        }
        if (vt == VAL_F32) {
            // entry q = entry j of group q/8: half j/4 of the values
            const uint64_t vq = (q % 8 / 4) * (n8u / 2) + q / 8 * 4 + q % 4;
            reinterpret_cast<float *>(tdata + 6 * base)[vq] = val[k];
        } else {
            reinterpret_cast<uint16_t *>(tdata + 4 * base)[q] = qval[k];
        }
        ctmp[off + p] = (uint16_t)(c & ((1u << wshift) - 1u));
        rtmp[off + p] = (uint16_t)i;
    }
}

// One work-item per group of 8 entries (entry offset 8*q): finds its tile,
// packs the group word (layout at the top of this file) into its unit.
// Padding (rtmp = kPadRow) repeats the previous key with column W.  bit 4
// of *bad: a step above kMaxRowStep, a key not below R or (quads) a quad
// holding two rows or starting with padding outside a tile's end padding
// (the layout rules all of them out; the SpMV kernels rely on it).
template <int W, bool Q>
__global__ __launch_bounds__(kWG) void tile_pack_kernel(const uint16_t *__restrict__ ctmp,
                                                        const uint16_t *__restrict__ rtmp,
                                                        const uint64_t *__restrict__ tile_off, uint64_t ntiles,
                                                        uint64_t nseg, uint32_t R, uint32_t vb,
                                                        uint8_t *__restrict__ tdata,
                                                        uint32_t *__restrict__ bad) {
    constexpr int CB = col_bits<W>();
    const uint64_t q = blockIdx.x * (uint64_t)kWG + threadIdx.x;
    const uint64_t L = 8 * q;
    if (L >= tile_off[ntiles]) return;
    // Last tile t with tile_off[t] <= L (empty tiles share their successor's
    // offset and come first, so this is the non-empty tile holding L).
    uint64_t lo = 0, hi = ntiles;  // tile_off[lo] <= L < tile_off[hi]
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) / 2;
        if (tile_off[mid] <= L)
            lo = mid;
        else
            hi = mid;
    }
    uint64_t base, e, n8u;  // as in tile_scatter_kernel
    if constexpr (Q) {
        const uint64_t b = lo / nseg;
        base = tile_off[b * nseg];
        e = L - base;
        n8u = tile_off[(b + 1) * nseg] - base;
    } else {
        const uint64_t off = tile_off[lo], p = L - off;
        base = off + p / kUE * kUE;
        e = p % kUE;
        n8u = std::min<uint64_t>(kUE, tile_off[lo + 1] - base);
    }

    const u16x8 r = *reinterpret_cast<const u16x8 *>(rtmp + L);
    const u16x8 c = *reinterpret_cast<const u16x8 *>(ctmp + L);
    uint32_t w0, big;
    uint32_t cols[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) cols[j] = r[j] == kPadRow ? (uint32_t)W : c[j];
    if constexpr (Q) {
        uint32_t kA = r[0], kB = r[4] == kPadRow ? kA : r[4];
        if (kA == kPadRow) {
            // A whole group of the padding at a tile's end: the key of the
            // tile's last row (its last quad starts with an entry), so the
            // padding extends that row.
            uint64_t pq = L;
            while (pq > tile_off[lo] && rtmp[pq - 4] == kPadRow) pq -= 4;
            if (pq > tile_off[lo]) kA = kB = rtmp[pq - 4];
        }
        big = (kA == kPadRow || kA >= R || kB >= R || kB < kA) ? 1u : 0u;
#pragma unroll
        for (int j = 1; j < 8; ++j) big |= (r[j] != kPadRow && r[j] != (j < 4 ? kA : kB)) ? 1u : 0u;
        w0 = kA | (kB << 16);
    } else {
        uint32_t prev = r[0] == kPadRow ? 0u : r[0];
        w0 = prev;
        big = prev >= R ? 1u : 0u;
#pragma unroll
        for (int j = 1; j < 8; ++j) {
            const uint32_t v = r[j];
            const uint32_t d = v == kPadRow ? 0u : v - prev;
            prev = v == kPadRow ? prev : v;
            big |= (d > kMaxRowStep || prev >= R) ? 1u : 0u;
            w0 |= (d & 3u) << (14 + 2 * j);
        }
    }
    uint64_t lo64 = 0;
    uint32_t w3 = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int bit = j * CB;
        const uint64_t cv = cols[j];
        if (bit < 64) lo64 |= cv << bit;
        if (bit + CB > 64) w3 |= (uint32_t)(bit >= 64 ? cv << (bit - 64) : cv >> (64 - bit));
    }
    u32x4 word;
    word.x = w0;
    word.y = (uint32_t)lo64;
    word.z = (uint32_t)(lo64 >> 32);
    word.w = w3;
    *reinterpret_cast<u32x4 *>(tdata + (vb + 2) * base + vb * n8u + 2 * e) = word;
    if (big) atomicOr(bad, 4u);
}

// val[k] <- its 16-bit code as f32 (without the row's scale), in place: the
// values the transposed tiles store (bit 8 of *bad as in quantize_rows_kernel).
__global__ void quantize_in_place_kernel(const uint64_t *__restrict__ rowptr, float *__restrict__ val, uint32_t rows,
                                         const float *__restrict__ scale, uint32_t vt, uint32_t *__restrict__ bad) {
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    const float inv = 1.0f / scale[row];
    bool neg = false;
    for (uint64_t k = rowptr[row]; k < rowptr[row + 1]; ++k) {
        const float t = val[k] * inv;
        if (is_uf16(vt)) {
            neg = neg || t < 0.0f;
            val[k] = from_uf16(to_uf16(fminf(fmaxf(t, 0.0f), 1.0f), kUf16ExpBits), kUf16ExpBits);
        } else {
            val[k] = vt == VAL_BF16 ? from_bf16(to_bf16(t)) : __half2float(__float2half_rn(t));
        }
    }
    if (neg && bad) atomicOr(bad, 8u);
}

// codes[k] = the 16-bit value q[k] as f32 (without its row's scale).
__global__ void decode_codes_kernel(const uint16_t *__restrict__ q, uint64_t n, uint32_t vt, float *__restrict__ codes) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x)
        codes[k] = vt == VAL_BF16   ? from_bf16(q[k])
                   : is_uf16(vt) ? from_uf16(q[k], kUf16ExpBits)
                                 : __half2float(__ushort_as_half(q[k]));
}

__global__ void fill_u16_kernel(uint16_t *__restrict__ p, uint16_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }
constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

size_t lds_bytes(uint32_t W, uint32_t R, uint32_t span, uint32_t BC) {
    return ((size_t)W + 4 + R + (BC == 1 ? kWG : 0)) * BC * 4 + (size_t)span * 16;
}

// stream_spmv_kernel: two x slots, the accumulator, the carries.
size_t stream_lds_bytes(uint32_t W, uint32_t R, uint32_t BC) {
    return (2 * ((size_t)W + 4) + R) * BC * 4;
}

template <int W, int VT, int BC>
hipError_t launch_k(const TilePlan &p, const TiledMatrix &tm, const TileWork &wk, const float *x, float *part,
                    hipStream_t s) {
    hipLaunchKernelGGL((blocked_spmv_kernel<W, VT, BC, pipeline_depth<W, VT, BC>()>), dim3(wk.nwg), dim3(kWG),
                       lds_bytes(W, p.R, wk.span, BC), s,
                       p.R, wk.span, wk.ntl, wk.wl, wk.order, wk.tiles, tm.tdata, x, part);
    return hipGetLastError();
}

template <int W, int VT, int BC, int NT>
hipError_t launch_stream_nt(const TilePlan &p, const TiledMatrix &tm, const TileWork &wk, const float *x, float *part,
                            hipStream_t s) {
    if (!wk.order) return hipErrorInvalidValue;
    hipLaunchKernelGGL((stream_spmv_kernel<W, VT, BC, NT>), dim3(wk.nwg), dim3(NT), stream_lds_bytes(W, p.R, BC), s,
                       p.R, p.nseg, tm.tile_off, wk.wl, wk.order, wk.tiles, tm.tdata, x, part);
    return hipGetLastError();
}

template <int W, int VT, int BC>
hipError_t launch_stream(const TilePlan &p, const TiledMatrix &tm, const TileWork &wk, const float *x, float *part,
                         uint32_t nt, hipStream_t s) {
    switch (nt) {
        case 128: return launch_stream_nt<W, VT, BC, 128>(p, tm, wk, x, part, s);
        case 256: return launch_stream_nt<W, VT, BC, 256>(p, tm, wk, x, part, s);
        case 0:
        case 512: return launch_stream_nt<W, VT, BC, 512>(p, tm, wk, x, part, s);
        default: return hipErrorInvalidValue;
    }
}

template <int VT>
hipError_t launch_vt(const TilePlan &p, const TiledMatrix &tm, const TileWork &wk, const float *x, float *part,
                     uint32_t BC, uint32_t nt, hipStream_t s) {
    if (p.quad != (BC > 1)) return hipErrorInvalidValue;  // quads (block streams) exactly for several
    if (BC == 1 && p.W == 1024) return launch_k<1024, VT, 1>(p, tm, wk, x, part, s);
    if (BC == 1 && p.W == 2048) return launch_k<2048, VT, 1>(p, tm, wk, x, part, s);
    if (BC == 4 && p.W == 512) return launch_stream<512, VT, 4>(p, tm, wk, x, part, nt, s);
    if (BC == 4 && p.W == 1024) return launch_stream<1024, VT, 4>(p, tm, wk, x, part, nt, s);
    if (BC == 8 && p.W == 512) return launch_stream<512, VT, 8>(p, tm, wk, x, part, nt, s);
    return hipErrorInvalidValue;
}

size_t scan_bytes(const TilePlan &p) {
    size_t scan = 0;
    (void)rocprim::exclusive_scan(nullptr, scan, (const uint64_t *)nullptr, (uint64_t *)nullptr, uint64_t(0),
                                  (size_t)p.ntiles + 1, rocprim::plus<uint64_t>());
    return scan;
}

// Count-phase scratch: cnt and rowpos [rows*nseg], tile_padded, tile_off and
// tile_nz of the count, the scan's temporary storage, the flag word.
struct CountScratch {
    uint32_t *cnt, *rowpos;
    uint64_t *tile_padded, *tile_off;
    uint32_t *tile_nz;
    void *scan_tmp;
    size_t scan_size;
    uint32_t *bad;
};

CountScratch count_scratch(const TilePlan &p, void *scratch) {
    CountScratch c;
    char *q = static_cast<char *>(scratch);
    c.cnt = reinterpret_cast<uint32_t *>(q);
    q += align256(4ull * p.rows * p.nseg);
    c.rowpos = reinterpret_cast<uint32_t *>(q);
    q += align256(4ull * p.rows * p.nseg);
    c.tile_padded = reinterpret_cast<uint64_t *>(q);
    q += align256(8 * (p.ntiles + 1));
    c.tile_off = reinterpret_cast<uint64_t *>(q);
    q += align256(8 * (p.ntiles + 1));
    c.tile_nz = reinterpret_cast<uint32_t *>(q);
    q += align256(4 * p.ntiles);
    c.scan_size = scan_bytes(p);
    c.scan_tmp = q;
    q += align256(c.scan_size);
    c.bad = reinterpret_cast<uint32_t *>(q);
    return c;
}

}  // namespace

TilePlan tile_plan(uint32_t rows, uint32_t cols, uint64_t nnz, uint32_t W, uint32_t R, uint32_t vt, bool quad) {
    TilePlan p;
    p.vt = vt;
    p.quad = quad;
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

bool valid_tile_width(uint32_t W) { return W == 512 || W == 1024 || W == 2048; }
bool valid_tile_config(uint32_t W, uint32_t BC) {
    return (BC == 1 && (W == 1024 || W == 2048)) || (BC == 4 && (W == 512 || W == 1024)) || (BC == 8 && W == 512);
}
bool valid_tile_rows(uint32_t R) { return R == 512 || R == 1024 || R == 2048 || R == 4096 || R == 8192 || R == 16384; }

size_t tiled_lds_bytes(const TilePlan &p, uint32_t span, uint32_t BC) {
    return p.quad ? stream_lds_bytes(p.W, p.R, BC) : lds_bytes(p.W, p.R, span, BC);
}

size_t tiled_matrix_bytes(const TilePlan &p) {
    return align256(8 * (p.ntiles + 1)) + align256(4 * p.ntiles) + align256((p.vb() + 2) * p.cap) +
           align256(4 * (size_t)p.rows);
}

size_t tiled_part_bytes(const TilePlan &p, uint32_t nwg, uint32_t BC) {
    return sizeof(float) * (size_t)nwg * p.R * BC;
}

namespace {
// Block streams (quad tiles): windows of up to kSNT groups from at most two
// consecutive non-empty tiles of a row block, then runs of consecutive
// windows of nearly equal group counts, none leaving its row block.
TileWorkList stream_work_list(const TilePlan &p, const uint64_t *off, uint32_t target_wgs) {
    TileWorkList L;
    const uint64_t total = off[p.ntiles] / 8;
    const uint64_t per_run = std::max<uint64_t>(1, (total + std::max(1u, target_wgs) - 1) / std::max(1u, target_wgs));
    L.wl_first.assign((size_t)p.nblk + 1, 0);
    std::vector<uint64_t> gs;    // first group of each non-empty tile of the block, then the block's groups
    std::vector<uint32_t> segs;  // its segment
    std::vector<uint64_t> wg;    // groups of each window of the block
    for (uint32_t b = 0; b < p.nblk; ++b) {
        L.wl_first[b] = L.nwg;
        const uint64_t base = off[(uint64_t)b * p.nseg];
        gs.clear();
        segs.clear();
        for (uint32_t sg = 0; sg < p.nseg; ++sg) {
            const uint64_t t = (uint64_t)b * p.nseg + sg;
            if (off[t] % (8 * kQuadAlign)) L.bad = 1;
            if (off[t + 1] == off[t]) continue;
            gs.push_back((off[t] - base) / 8);
            segs.push_back(sg);
        }
        const uint64_t G = (off[(uint64_t)(b + 1) * p.nseg] - base) / 8;
        const size_t nt = segs.size();
        gs.push_back(G);
        const uint32_t first_win = L.ntl;
        wg.clear();
        size_t ta = 0;
        for (uint64_t pos = 0; pos < G;) {
            while (gs[ta + 1] <= pos) ++ta;  // the tile holding group pos
            const uint64_t end = std::min<uint64_t>(pos + kSNT, gs[std::min(ta + 2, nt)]);
            const uint64_t split = std::min(gs[ta + 1], end) - pos;
            const uint32_t sb = ta + 1 < nt ? segs[ta + 1] : segs[ta];
            L.tiles.insert(L.tiles.end(), {(uint32_t)pos, (uint32_t)((end - pos) | (split << 16) | ((ta & 1) << 31)),
                                           segs[ta], sb});
            ++L.ntl;
            wg.push_back(end - pos);
            pos = end;
        }
        // Runs of nearly equal group counts over the block's windows.
        uint64_t left = G, runs = (G + per_run - 1) / per_run;
        size_t wi = 0;
        while (wi < wg.size()) {
            const uint64_t want = (left + runs - 1) / runs;
            const size_t w0 = wi;
            uint64_t got = 0;
            while (wi < wg.size() && wi - w0 < kRunWindows && (got == 0 || got + wg[wi] / 2 <= want)) got += wg[wi++];
            L.wl.insert(L.wl.end(), {b, first_win + (uint32_t)w0, (uint32_t)(wi - w0), 0u});
            ++L.nwg;
            left -= got;
            runs = runs > 1 ? runs - 1 : 1;
        }
    }
    L.wl_first[p.nblk] = L.nwg;
    // Launch order: by the column segment each run starts in, then by row
    // block, so that the workgroups in flight together stage the x segments
    // of a band of columns, which then stay in L2 (in row-block order every
    // row block re-reads all of x from HBM).  A run's place in wl, and so its
    // partial sums and their order in the reduction, do not change.
    L.order.resize(L.nwg);
    for (uint32_t r = 0; r < L.nwg; ++r) L.order[r] = r;
    auto first_seg = [&](uint32_t r) { return L.tiles[4 * (size_t)L.wl[4 * (size_t)r + 1] + 2]; };
    std::stable_sort(L.order.begin(), L.order.end(),
                     [&](uint32_t a, uint32_t b) { return first_seg(a) < first_seg(b); });
    return L;
}
}  // namespace

TileWorkList tile_work_list(const TilePlan &p, const uint32_t *nz, const uint64_t *off, uint32_t target_wgs) {
    if (p.quad) return stream_work_list(p, off, target_wgs);
    const uint64_t UE = kUE;
    TileWorkList L;
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
    // Launch order: row-block order (workgroup w takes run w).  With
    // MLEM_TILE_ORDER=cols the runs go by the column segment of their first
    // tile, then by row block, as for the block streams, so that the
    // workgroups in flight together stage the x segments of a band of
    // columns: on the 3D scanner matrix (1.53e9 nonzeros) that fetches 0.7%
    // fewer bytes per launch (9.53 GB instead of 9.59), but each launch takes
    // 1.5-3% longer (the streams in flight are spread over the whole matrix),
    // so row-block order stays the default.  The partial sums and the images
    // are the same either way.
    const char *ord = std::getenv("MLEM_TILE_ORDER");
    if (ord && std::strcmp(ord, "cols") == 0) {
        L.order.resize(L.nwg);
        for (uint32_t r = 0; r < L.nwg; ++r) L.order[r] = r;
        auto first_seg = [&](uint32_t r) { return L.tiles[4 * (size_t)L.wl[4 * (size_t)r + 1]]; };
        std::stable_sort(L.order.begin(), L.order.end(),
                         [&](uint32_t a, uint32_t b) { return first_seg(a) < first_seg(b); });
    }
    return L;
}

size_t tile_count_scratch_bytes(const TilePlan &p) {
    return 2 * align256(4ull * p.rows * p.nseg) + 2 * align256(8 * (p.ntiles + 1)) + align256(4 * p.ntiles) +
           align256(scan_bytes(p)) + 256;
}

size_t tile_fill_scratch_bytes(const TilePlan &p) {
    return 2 * align256(2 * p.cap) + (p.vt == VAL_F32 ? 0 : align256(2 * p.nnz));
}

TiledMatrix tiled_matrix_view(const TilePlan &p, void *mem) {
    TiledMatrix t;
    char *c = static_cast<char *>(mem);
    t.tile_off = reinterpret_cast<uint64_t *>(c);
    c += align256(8 * (p.ntiles + 1));
    t.tile_nz = reinterpret_cast<uint32_t *>(c);
    c += align256(4 * p.ntiles);
    t.tdata = reinterpret_cast<uint8_t *>(c);
    c += align256((p.vb() + 2) * p.cap);
    t.scale = p.vt == VAL_F32 ? nullptr : reinterpret_cast<float *>(c);
    return t;
}

hipError_t launch_tile_count(TilePlan &p, const uint64_t *rowptr, const uint32_t *col, void *scratch,
                             uint32_t *flags_host, hipStream_t s) {
    CountScratch c = count_scratch(p, scratch);
    uint32_t wshift = 0;
    while ((1u << wshift) < p.W) ++wshift;
    hipError_t e;
    *flags_host = 0;
    if ((e = hipMemsetAsync(c.bad, 0, sizeof(uint32_t), s)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(c.cnt, 0, 4ull * p.rows * p.nseg, s)) != hipSuccess) return e;
    if (p.rows) {
        hipLaunchKernelGGL(rows_sorted_kernel, dim3(blocks_for(p.rows, kWG / 64)), dim3(kWG), 0, s, rowptr, col,
                           p.rows, c.bad);
        hipLaunchKernelGGL(tile_count_kernel, dim3(blocks_for(p.rows, kWG / 64)), dim3(kWG), 0, s, rowptr, col,
                           p.rows, p.nseg, wshift, c.cnt);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // The layout is only correct (and the fill only in bounds) for sorted
    // rows: check the flag first.
    if ((e = hipMemcpyAsync(flags_host, c.bad, sizeof(uint32_t), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    if (*flags_host) return hipSuccess;
    if (p.ntiles) {
        hipLaunchKernelGGL(tile_rowpos_kernel, dim3(blocks_for(p.ntiles, kWG)), dim3(kWG), 0, s, c.cnt, p.rows, p.R,
                           p.nseg, p.ntiles, p.quad, c.rowpos, c.tile_nz, c.tile_padded);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if ((e = hipMemsetAsync(c.tile_padded + p.ntiles, 0, sizeof(uint64_t), s)) != hipSuccess) return e;
    size_t sb = c.scan_size;
    if ((e = rocprim::exclusive_scan(c.scan_tmp, sb, c.tile_padded, c.tile_off, uint64_t(0), (size_t)p.ntiles + 1,
                                     rocprim::plus<uint64_t>(), s)) != hipSuccess)
        return e;
    uint64_t total = 0;
    if ((e = hipMemcpyAsync(&total, c.tile_off + p.ntiles, sizeof(uint64_t), hipMemcpyDeviceToHost, s)) != hipSuccess)
        return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    p.cap = total + 8;  // a multiple of 8, at least one group
    return hipSuccess;
}

const uint64_t *tile_count_offsets(const TilePlan &p, void *scratch) { return count_scratch(p, scratch).tile_off; }
const uint32_t *tile_count_sizes(const TilePlan &p, void *scratch) { return count_scratch(p, scratch).tile_nz; }

hipError_t launch_tile_fill(const TilePlan &p, const uint64_t *rowptr, const uint32_t *col, const float *val,
                            void *mem, void *scratch, void *fill_scratch, uint32_t *flags_host, hipStream_t s) {
    return launch_tile_fill_into(p, rowptr, col, val, tiled_matrix_view(p, mem), scratch, fill_scratch, flags_host,
                                 s);
}

hipError_t launch_tile_fill_into(const TilePlan &p, const uint64_t *rowptr, const uint32_t *col, const float *val,
                                 const TiledMatrix &tm, void *scratch, void *fill_scratch, uint32_t *flags_host,
                                 hipStream_t s) {
    CountScratch c = count_scratch(p, scratch);
    uint16_t *ctmp = static_cast<uint16_t *>(fill_scratch);
    uint16_t *rtmp = reinterpret_cast<uint16_t *>(static_cast<char *>(fill_scratch) + align256(2 * p.cap));
    uint32_t wshift = 0;
    while ((1u << wshift) < p.W) ++wshift;
    hipError_t e;
    *flags_host = 0;
    if ((e = hipMemcpyAsync(tm.tile_off, c.tile_off, 8 * (p.ntiles + 1), hipMemcpyDeviceToDevice, s)) != hipSuccess)
        return e;
    if (p.ntiles &&
        (e = hipMemcpyAsync(tm.tile_nz, c.tile_nz, 4 * p.ntiles, hipMemcpyDeviceToDevice, s)) != hipSuccess)
        return e;
    // Padding entries: value 0 (and, in the packing, column W).
    if ((e = hipMemsetAsync(tm.tdata, 0, (p.vb() + 2) * p.cap, s)) != hipSuccess) return e;
    uint16_t *qval = nullptr;
    if (tm.scale && p.rows) {
        qval = reinterpret_cast<uint16_t *>(static_cast<char *>(fill_scratch) + 2 * align256(2 * p.cap));
        if (p.unit_scale) {  // the values are 16-bit codes already: stored as they are
            if ((e = launch_fill(tm.scale, 1.0f, p.rows, s)) != hipSuccess) return e;
        } else {
            hipLaunchKernelGGL(row_absmax_kernel, dim3(blocks_for(p.rows, kWG / 64)), dim3(kWG), 0, s, rowptr, val,
                               p.rows, value_scale_mult(p.vt), tm.scale);
        }
        hipLaunchKernelGGL(quantize_rows_kernel, dim3(blocks_for(p.rows, 64)), dim3(64), 0, s, rowptr, val, p.rows,
                           tm.scale, p.vt, qval, c.bad);
    }
    if ((e = hipMemsetAsync(rtmp, 0xff, 2 * p.cap, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(fill_u16_kernel, dim3(1024), dim3(kWG), 0, s, ctmp, (uint16_t)p.W, p.cap);
    if (p.rows) {
        hipLaunchKernelGGL(row_segscan_kernel, dim3(blocks_for(p.rows, kWG)), dim3(kWG), 0, s, c.cnt, p.rows, p.nseg);
        hipLaunchKernelGGL(tile_scatter_kernel, dim3(blocks_for(p.rows, kWG / 64)), dim3(kWG), 0, s, rowptr, col, val,
                           p.rows, p.R, p.nseg, wshift, c.cnt, c.rowpos, tm.tile_off, p.vt, p.quad, qval, tm.tdata,
                           ctmp, rtmp);
    }
    if (p.ntiles) {
        const dim3 grid(blocks_for(p.cap / 8, kWG));
#define MLEM_PACK(WW, QQ)                                                                                       \
    hipLaunchKernelGGL((tile_pack_kernel<WW, QQ>), grid, dim3(kWG), 0, s, ctmp, rtmp, tm.tile_off, p.ntiles, p.nseg, p.R, \
                       p.vb(), tm.tdata, c.bad)
        if (p.W == 512 && p.quad)
            MLEM_PACK(512, true);
        else if (p.W == 512)
            MLEM_PACK(512, false);
        else if (p.W == 1024 && p.quad)
            MLEM_PACK(1024, true);
        else if (p.W == 1024)
            MLEM_PACK(1024, false);
        else if (p.quad)
            MLEM_PACK(2048, true);
        else
            MLEM_PACK(2048, false);
#undef MLEM_PACK
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(flags_host, c.bad, sizeof(uint32_t), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

hipError_t launch_tiled_partials(const TilePlan &p, const TiledMatrix &tm, const TileWork &wk, const float *x,
                                 float *part, uint32_t BC, uint32_t nt, hipStream_t s) {
    if (p.rows == 0 || wk.nwg == 0) return hipSuccess;
    switch (p.vt) {
        case VAL_F32: return launch_vt<VAL_F32>(p, tm, wk, x, part, BC, nt, s);
        case VAL_BF16: return launch_vt<VAL_BF16>(p, tm, wk, x, part, BC, nt, s);
        case VAL_UF16: return launch_vt<VAL_UF16>(p, tm, wk, x, part, BC, nt, s);
        default: return launch_vt<VAL_F16>(p, tm, wk, x, part, BC, nt, s);
    }
}

bool valid_stream_threads(uint32_t nt) { return nt == 0 || nt == 128 || nt == 256 || nt == 512; }

namespace {
template <int EPI, int BC>
void launch_reduce_epi(const TilePlan &p, const TileWork &wk, const float *part, const float *scale, uint32_t rshift,
                       const EpiArgs &ea, hipStream_t s) {
    if (wk.max_runs >= kSlicedRuns)
        hipLaunchKernelGGL((blocked_reduce_sliced_kernel<EPI, BC>), dim3(blocks_for(p.rows, 64)), dim3(64, kSlices), 0,
                           s, part, wk.wl_first, rshift, p.rows, scale, ea);
    else
        hipLaunchKernelGGL((blocked_reduce_kernel<EPI, BC>), dim3(blocks_for(p.rows, kWG)), dim3(kWG), 0, s, part,
                           wk.wl_first, rshift, p.rows, scale, ea);
}

template <int BC>
hipError_t launch_reduce_bc(const TilePlan &p, const TileWork &wk, const float *part, const float *scale, Epi epi,
                            const EpiArgs &ea, hipStream_t s) {
    uint32_t rshift = 0;
    while ((1u << rshift) < p.R) ++rshift;
    switch (epi) {
        case EPI_PLAIN: launch_reduce_epi<EPI_PLAIN, BC>(p, wk, part, scale, rshift, ea, s); break;
        case EPI_RATIO: launch_reduce_epi<EPI_RATIO, BC>(p, wk, part, scale, rshift, ea, s); break;
        default: launch_reduce_epi<EPI_UPDATE, BC>(p, wk, part, scale, rshift, ea, s); break;
    }
    return hipGetLastError();
}
}  // namespace

hipError_t launch_tiled_reduce(const TilePlan &p, const TiledMatrix &tm, const TileWork &wk, const float *part,
                               uint32_t BC, Epi epi, const EpiArgs &ea, hipStream_t s) {
    if (p.rows == 0) return hipSuccess;
    switch (BC) {
        case 1: return launch_reduce_bc<1>(p, wk, part, tm.scale, epi, ea, s);
        case 4: return launch_reduce_bc<4>(p, wk, part, tm.scale, epi, ea, s);
        case 8: return launch_reduce_bc<8>(p, wk, part, tm.scale, epi, ea, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_quantize_in_place(const uint64_t *rowptr, float *val, uint32_t rows, uint32_t vt, float *scale,
                                    uint32_t *bad, hipStream_t s) {
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(row_absmax_kernel, dim3(blocks_for(rows, kWG / 64)), dim3(kWG), 0, s, rowptr, val, rows,
                       value_scale_mult(vt), scale);
    hipLaunchKernelGGL(quantize_in_place_kernel, dim3(blocks_for(rows, 64)), dim3(64), 0, s, rowptr, val, rows, scale,
                       vt, bad);
    return hipGetLastError();
}

hipError_t launch_quantize_codes(const uint64_t *rowptr, const float *val, uint32_t rows, uint64_t nnz, uint32_t vt,
                                 float *scale, float *codes, uint16_t *q, hipStream_t s) {
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(row_absmax_kernel, dim3(blocks_for(rows, kWG / 64)), dim3(kWG), 0, s, rowptr, val, rows,
                       value_scale_mult(vt), scale);
    hipLaunchKernelGGL(quantize_rows_kernel, dim3(blocks_for(rows, 64)), dim3(64), 0, s, rowptr, val, rows, scale, vt,
                       q, nullptr);
    if (nnz) {
        const uint64_t blocks = std::min<uint64_t>((nnz + kWG - 1) / kWG, 65536);
        hipLaunchKernelGGL(decode_codes_kernel, dim3((uint32_t)blocks), dim3(kWG), 0, s, q, nnz, vt, codes);
    }
    return hipGetLastError();
}

}  // namespace mlem
