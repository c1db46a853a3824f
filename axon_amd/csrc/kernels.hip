// HIP/CDNA4 (gfx950) kernels of the AXON engine.
//
// Geometry (shared with the reference so the synapse store is bit-identical,
// mm:30-43): a column is 32 neurons, each neuron has 32 dendritic branches of
// 256 presynaptic bits.  A (column, branch) tile of one plane is 64 uint4 =
// 1 KiB laid out [j][neuron] (mm:170), so one wave64 reads a whole tile with a
// single global_load_dwordx4 per lane: lane L = j*32 + neuron, j selecting the
// low or high 128 presynaptic bits.  The two halves of the wave are combined
// with one cross-half exchange per branch.
//
// All arithmetic is u32 integer with the reference's wrap-around semantics.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <mutex>
#include <type_traits>

#include "engine.h"

namespace axon {

namespace {

__device__ __forceinline__ uint32_t hash32(uint32_t x) {  // mm:54
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }
__device__ __forceinline__ uint32_t ballot32(bool p) { return (uint32_t)__ballot(p); }
__device__ __forceinline__ uint32_t popc(uint32_t x) { return (uint32_t)__popc(x); }

// Sum over all 64 lanes (u32 wrap-around), result in every lane.
__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
    x += __shfl_xor(x, 32);
    x += __shfl_xor(x, 16);
    x += __shfl_xor(x, 8);
    x += __shfl_xor(x, 4);
    x += __shfl_xor(x, 2);
    x += __shfl_xor(x, 1);
    return x;
}
// Inclusive prefix sum over the 32 lanes of each half (u32 wrap-around):
// row_shr 1, 2, 4, 8 within the 16-lane rows, then row 0's (row 2's) total
// broadcast from lane 15 (47) into row 1 (row 3) only.  Lane 31 (63) holds
// the half's total.
__device__ __forceinline__ uint32_t half_scan(uint32_t x) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);
    return x;
}
// 32x32 bit transpose within each half-wave: lane s of a half returns, as bit
// n, bit s of lane n's x (the same half).  Five butterfly stages, each
// exchanging blocks with the lane j apart (ds_swizzle, xor mask j).
__device__ __forceinline__ uint32_t half_transpose32(uint32_t x) {
    const uint32_t n = lane_id() & 31u;
    uint32_t y;
    y = (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x401F);
    x = (n & 16u) ? ((x & 0xFFFF0000u) | (y >> 16)) : ((x & 0x0000FFFFu) | (y << 16));
    y = (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x201F);
    x = (n & 8u) ? ((x & 0xFF00FF00u) | ((y >> 8) & 0x00FF00FFu)) : ((x & 0x00FF00FFu) | ((y & 0x00FF00FFu) << 8));
    y = (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x101F);
    x = (n & 4u) ? ((x & 0xF0F0F0F0u) | ((y >> 4) & 0x0F0F0F0Fu)) : ((x & 0x0F0F0F0Fu) | ((y & 0x0F0F0F0Fu) << 4));
    y = (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x081F);
    x = (n & 2u) ? ((x & 0xCCCCCCCCu) | ((y >> 2) & 0x33333333u)) : ((x & 0x33333333u) | ((y & 0x33333333u) << 2));
    y = (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x041F);
    x = (n & 1u) ? ((x & 0xAAAAAAAAu) | ((y >> 1) & 0x55555555u)) : ((x & 0x55555555u) | ((y & 0x55555555u) << 1));
    return x;
}
// Inclusive prefix sum over all 64 lanes: half_scan, then the first half's
// total (lane 31) broadcast into rows 2 and 3.
__device__ __forceinline__ uint32_t wave_scan(uint32_t x) {
    x = half_scan(x);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);
    return x;
}
// Sum over the 32 lanes of each half.
__device__ __forceinline__ uint32_t half_sum(uint32_t x) {
    x += __shfl_xor(x, 16);
    x += __shfl_xor(x, 8);
    x += __shfl_xor(x, 4);
    x += __shfl_xor(x, 2);
    x += __shfl_xor(x, 1);
    return x;
}
__device__ __forceinline__ uint32_t half_or(uint32_t x) {
    x |= __shfl_xor(x, 16);
    x |= __shfl_xor(x, 8);
    x |= __shfl_xor(x, 4);
    x |= __shfl_xor(x, 2);
    x |= __shfl_xor(x, 1);
    return x;
}

// Learning-column lists of the side-statistics path: per step parity, NLISTS
// lists, so the list-slot atomics of different columns hit different
// counters; each counter sits on its own 128-byte line.  Column c goes to
// list (c / 4) % 8.  A list holds at most list_cols(ncol) columns.
__device__ __forceinline__ uint32_t* list_count(const NetDev& d, uint32_t par, uint32_t l) {
    return d.learnCnt + (par * NLISTS + l) * 32u;
}
__device__ __forceinline__ uint32_t list_of(uint32_t col) { return (col >> 2) & (NLISTS - 1u); }

// Word-major tile addressing (see TILE_WORDS): word w of neuron n of a tile.
__device__ __forceinline__ size_t word_idx(uint32_t tile, uint32_t w, uint32_t n) {
    return ((size_t)tile * 8u + w) * 32u + n;
}
// Range-checked buffer access: a lane whose byte offset is at or past `bytes`
// reads 0 and writes nothing, and issues no memory request.  Used to load or
// store only the words a lane needs without branching around the access.
constexpr uint32_t SKIP = 0x80000000u;  // an offset past every range used here
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ uint32_t buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
}
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t v) {
    __builtin_amdgcn_raw_buffer_store_b32(v, r, off, 0, 0);
}
__device__ __forceinline__ void buf_st16(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t v) {
    __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v, r, off, 0, 0);
}

// 8-bit mask of the nonzero bytes of x (bit i = byte i).
__device__ __forceinline__ uint32_t byte_mask(uint64_t x) {
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) m |= (((x >> (8 * i)) & 0xFFull) != 0ull ? 1u : 0u) << i;
    return m;
}

// The 8 spike words of presynaptic group g (columns 8g..8g+7) as this lane's
// half: words 0-3 in lanes 0-31, words 4-7 in lanes 32-63.  A vector load,
// so a wave keeps many of them in flight (scalar loads of the same words are
// waited for one at a time).
__device__ __forceinline__ uint4 group_words(const uint32_t* __restrict__ spk, uint32_t g, bool hi) {
    return reinterpret_cast<const uint4*>(spk)[2u * g + (hi ? 1u : 0u)];
}


// ---- exact k-of-32 winners-take-all with margin (kwta_ballot_margin, mm:134-144) ----
// Lanes 0..31 hold the column's 32 values (lanes 32..63 duplicate them).
// The reference ranks by (value desc, lane asc): rank_l = #{j: v_j > v_l or
// (v_j == v_l and j < l)}; winners are rank < k; vk / vk1 are the values at
// rank k-1 / k (0 when that rank does not exist).  Here the m-th largest value
// is found by a bitwise radix select over ballots, which gives the same sets:
//   winners = {v > vk} plus the (k - #{v > vk}) lowest lanes with v == vk.
// h selects the 32-lane half that holds the column (its ballot bits).
__device__ __forceinline__ uint32_t hball(bool p, uint32_t h) { return (uint32_t)(__ballot(p) >> (32u * h)); }
__device__ __forceinline__ uint32_t kth_largest(uint32_t v, uint32_t m, int hb, uint32_t h) {
    uint32_t x = 0;
    for (int b = hb; b >= 0; --b) {
        uint32_t t = x | (1u << b);
        if (popc(hball(v >= t, h)) >= m) x = t;
    }
    return x;
}
// Maximum over the 32 lanes of each half, in every lane of that half.
__device__ __forceinline__ uint32_t half_max(uint32_t x) {
    x = max(x, (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x401F));
    x = max(x, (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x201F));
    x = max(x, (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x101F));
    x = max(x, (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x081F));
    x = max(x, (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x041F));
    return x;
}
struct Kwta { uint32_t ball, vk, vk1; };
__device__ __forceinline__ Kwta kwta32(uint32_t v, uint32_t k, uint32_t n, uint32_t h = 0u) {
    const uint32_t ov = half_or(v);
    const int hb = ov ? 31 - __clz(ov) : -1;
    Kwta r;
    r.vk = (k >= 1u && k <= 32u) ? kth_largest(v, k, hb, h) : 0u;
    // the (k+1)-th largest: vk itself when more than k values reach it, else the
    // largest value below vk (k = 0: the largest value)
    if (k == 0u) r.vk1 = half_max(v);
    else if (k <= 31u) r.vk1 = popc(hball(v >= r.vk, h)) > k ? r.vk : half_max(v < r.vk ? v : 0u);
    else r.vk1 = 0u;
    if (k == 0u) r.ball = 0u;
    else if (k >= 32u) r.ball = 0xFFFFFFFFu;
    else {
        const uint32_t gt = hball(v > r.vk, h);
        const uint32_t eq = hball(v == r.vk, h);
        const uint32_t need = k - popc(gt);
        const bool win = (v > r.vk) || (v == r.vk && popc(eq & ((1u << n) - 1u)) < need);
        r.ball = hball(win, h);
    }
    return r;
}

// ---- bit-serial 4-plane counters on one 32-synapse word (mm:371-420) ----
__device__ __forceinline__ void bs_inc(uint32_t& a0, uint32_t& a1, uint32_t& a2, uint32_t& a3, uint32_t m) {
    uint32_t c = m, t;
    t = a0 ^ c; c = a0 & c; a0 = t;
    t = a1 ^ c; c = a1 & c; a1 = t;
    t = a2 ^ c; c = a2 & c; a2 = t;
    t = a3 ^ c; c = a3 & c; a3 = t;
    a0 |= c; a1 |= c; a2 |= c; a3 |= c;
}
__device__ __forceinline__ void bs_dec(uint32_t& a0, uint32_t& a1, uint32_t& a2, uint32_t& a3, uint32_t m) {
    uint32_t b = m, t;
    t = a0 ^ b; b = (~a0) & b; a0 = t;
    t = a1 ^ b; b = (~a1) & b; a1 = t;
    t = a2 ^ b; b = (~a2) & b; a2 = t;
    t = a3 ^ b; b = (~a3) & b; a3 = t;
    const uint32_t keep = ~b;
    a0 &= keep; a1 &= keep; a2 &= keep; a3 &= keep;
}
__device__ __forceinline__ uint32_t bs_lt(uint32_t a3, uint32_t a2, uint32_t a1, uint32_t a0, uint32_t K) {
    uint32_t lt = 0, eq = 0xFFFFFFFFu, kb;
    kb = (K >> 3) & 1u; if (kb) lt |= eq & ~a3; eq &= kb ? a3 : ~a3;
    kb = (K >> 2) & 1u; if (kb) lt |= eq & ~a2; eq &= kb ? a2 : ~a2;
    kb = (K >> 1) & 1u; if (kb) lt |= eq & ~a1; eq &= kb ? a1 : ~a1;
    kb = (K >> 0) & 1u; if (kb) lt |= eq & ~a0;
    return lt;
}
__device__ __forceinline__ uint32_t bs_gt(uint32_t a3, uint32_t a2, uint32_t a1, uint32_t a0, uint32_t K) {
    uint32_t gt = 0, eq = 0xFFFFFFFFu, kb;
    kb = (K >> 3) & 1u; if (!kb) gt |= eq & a3; eq &= kb ? a3 : ~a3;
    kb = (K >> 2) & 1u; if (!kb) gt |= eq & a2; eq &= kb ? a2 : ~a2;
    kb = (K >> 1) & 1u; if (!kb) gt |= eq & a1; eq &= kb ? a1 : ~a1;
    kb = (K >> 0) & 1u; if (!kb) gt |= eq & a0;
    return gt;
}
__device__ __forceinline__ void bs_set(uint32_t& a0, uint32_t& a1, uint32_t& a2, uint32_t& a3, uint32_t m,
                                       uint32_t C) {
    const uint32_t keep = ~m;
    a0 = (a0 & keep) | (((C >> 0) & 1u) ? m : 0u);
    a1 = (a1 & keep) | (((C >> 1) & 1u) ? m : 0u);
    a2 = (a2 & keep) | (((C >> 2) & 1u) ? m : 0u);
    a3 = (a3 & keep) | (((C >> 3) & 1u) ? m : 0u);
}
__device__ __forceinline__ void bs_clip(uint32_t& a0, uint32_t& a1, uint32_t& a2, uint32_t& a3, uint32_t lo,
                                        uint32_t hi) {
    const uint32_t l = bs_lt(a3, a2, a1, a0, lo);
    const uint32_t g = bs_gt(a3, a2, a1, a0, hi);
    bs_set(a0, a1, a2, a3, l, lo);
    bs_set(a0, a1, a2, a3, g, hi);
}

// One 32-synapse word of k_learn's per-tile update (mm:471-487).
__device__ __forceinline__ void learn_word(uint32_t& a0, uint32_t& a1, uint32_t& a2, uint32_t& a3, uint32_t sp,
                                           bool want, const Params& P, uint32_t& nLTP, uint32_t& nLTD) {
    const uint32_t co = want ? sp : 0u;
    const uint32_t nco = popc(co);
    if (P.phase == 1u) {
        bs_inc(a0, a1, a2, a3, co);
        nLTP += nco;
    } else if (P.phase == 2u) {
        for (uint32_t rr = 0; rr < P.ltdRep; ++rr) bs_dec(a0, a1, a2, a3, co);
        nLTD += nco * P.ltdRep;
    } else {
        bs_inc(a0, a1, a2, a3, co);
        bs_dec(a0, a1, a2, a3, (want && P.depress != 0u) ? ~sp : 0u);
    }
    if (P.hetero != 0u && P.phase == 1u) bs_dec(a0, a1, a2, a3, want ? ~sp : 0u);
    if (P.bandHi != 0u) bs_clip(a0, a1, a2, a3, P.bandLo, P.bandHi);
}

// ============================================================== init
// k_init_w, mm:86-104.  Thread gg produces the values of the reference's
// uint4 number gg (tile gg / 64, half (gg / 32) & 1, neuron gg & 31) and
// stores its four words at their word-major places.
__global__ __launch_bounds__(256) void k_init_w(uint32_t* __restrict__ W0, uint32_t* __restrict__ W1,
                                                uint32_t* __restrict__ W2, uint32_t* __restrict__ W3,
                                                uint32_t nAnd, uint32_t seed, uint32_t n) {
    const uint32_t gg = blockIdx.x * 256u + threadIdx.x;
    if (gg >= n) return;
    const uint32_t g = gg ^ (seed * 0x9E3779B9u);
    uint4 m = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
    for (uint32_t i = 0; i < nAnd; ++i) {
        m.x &= hash32(g * 32u + i * 4u + 0u);
        m.y &= hash32(g * 32u + i * 4u + 1u);
        m.z &= hash32(g * 32u + i * 4u + 2u);
        m.w &= hash32(g * 32u + i * 4u + 3u);
    }
    const uint4 c = make_uint4(hash32(g * 8u + 11u), hash32(g * 8u + 13u), hash32(g * 8u + 17u),
                               hash32(g * 8u + 19u));
    const uint32_t mw[4] = {m.x, m.y, m.z, m.w}, cw[4] = {c.x, c.y, c.z, c.w};
    const uint32_t tile = gg >> 6, h = (gg >> 5) & 1u, nn = gg & 31u;
#pragma unroll
    for (uint32_t k = 0; k < 4u; ++k) {
        const size_t i = word_idx(tile, 4u * h + k, nn);
        W3[i] = mw[k];
        W2[i] = ~mw[k] | cw[k];
        W1[i] = cw[k];
        W0[i] = ~cw[k];
    }
}

// k_init_conn, mm:109-117
__global__ __launch_bounds__(256) void k_init_conn(uint32_t* __restrict__ conn, uint32_t nGrp, uint32_t nInGrp,
                                                   uint32_t inBr, uint32_t seed, uint32_t n) {
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    if (g >= n) return;
    const uint32_t s = g % NSLOT;
    const uint32_t h = hash32(g * 2246822519u + 7u + seed * 2654435761u);
    conn[g] = (s < inBr) ? (h % nInGrp) : (nInGrp + h % (nGrp - nInGrp));
}

__global__ __launch_bounds__(256) void k_fill(uint32_t* __restrict__ p, uint32_t v, size_t n) {
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256ull) p[i] = v;
}

// Start of a presentation: both spike buffers and the stats words to zero
// (k_clear mm:497, memset of bStat mm:824).  A kernel, not a memset, so the
// whole presentation can be captured in one hipGraph.
__global__ __launch_bounds__(256) void k_clear(uint32_t* __restrict__ a, uint32_t* __restrict__ b, uint32_t ncol,
                                               uint32_t* __restrict__ st, uint32_t nst, uint8_t* __restrict__ nz,
                                               uint32_t* __restrict__ lcnt) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < ncol) { a[i] = 0u; b[i] = 0u; nz[i] = 0u; }
    if (i < 2u * NLISTS) lcnt[i * 32u] = 0u;  // the plasticity list counters (list_count)
    if (i < nst) st[i] = 0u;
}

// ============================================================== propagation
// k_drive, mm:150-190.  One wave per column; lane = (j, neuron).
// Branches whose presynaptic group carries no spike are skipped wave-uniformly
// (mm:169); the remaining tiles are fetched eight at a time so several 1 KiB
// loads are in flight per wave.  Spike words of a group are wave-uniform and
// come through scalar loads.
// `gl` is conn[col*32 + (lane & 31)] (loaded by the caller so that the load
// overlaps the staging of `fired`).  `fired`, when not null, is the LDS copy
// of the per-column "spiked last step" bytes: a branch is active when any of
// the 8 columns of its presynaptic group fired, one 8-byte LDS read.
// The per-column inputs that do not depend on which branches are active,
// loaded unconditionally before anything waits for memory: the lane's branch
// group (conn), its neuron's bias word, the column's previous spike word and
// its two work counters (only this column's wave writes them, so they are
// updated by a plain read-modify-write instead of an atomic).
struct DriveIn { uint32_t gl, biasw, prev; uint2 cnt; };
__device__ __forceinline__ DriveIn drive_inputs(const NetDev& d, const uint32_t* __restrict__ spkIn, uint32_t col) {
    const uint32_t n = lane_id() & 31u;
    return DriveIn{d.conn[col * NSLOT + n], d.bias[col * NPC + n], spkIn[col],
                   reinterpret_cast<const uint2*>(d.cntProp)[col]};
}

// LDS item list capacity of a half-wave: the branches are split where the
// column's items halve, so a half holds at most (256 + 8) / 2 items.
constexpr uint32_t ITEMS_PER_HALF = (NSLOT * 8u + 8u) / 2u;
// Columns per propagation workgroup (k_drive), one wave each.  Measured on
// one 131072-neuron network: 4 columns 16.05k GSyn/s, 8 columns 16.3k, 16
// columns 16.4k (a heavy learning column's voted lines spread over more
// half-waves).
constexpr uint32_t DRIVE_COLS = 16;

// Work items are (branch, spiking presynaptic column) pairs, one 128-byte W3
// line each, listed in LDS with their load offsets ready-made (the drive
// kernels are bound by instruction issue and latency, not bytes):
//   bits 0-7   branch * 8 + word  (W3 word index within the column tile)
//   bit  9     pad: its W3 offset (64 KiB on) is past both drive forms' resources
//   bits 11-30 presynaptic column 8g + word (ncol <= 2^20)
//   bit  31    last item of its branch (a sign test)
// A half's list is padded to a whole round with DRIVE_PAD_ITEM (never last;
// its W3 word reads as 0 without a request, and so does its spike word for
// any column count but 2^20).
constexpr uint32_t DRIVE_PAD_ITEM = 0x200u | (0xFFFFFu << 11);
__device__ __forceinline__ uint32_t drive_item(uint32_t br, uint32_t w, uint32_t g, bool last) {
    return (br << 3) + w + (last ? 0x80000000u : 0u) + ((8u * g + w) << 11);
}
// The contribution of a branch whose count is c (mm:175-183): mode 0 the
// count, mode 1 the count plus boost at the NMDA threshold, otherwise 1 at
// the threshold.
template <int MODE>
__device__ __forceinline__ uint32_t branch_out(uint32_t c, bool nmda, uint32_t boost) {
    return MODE == 0 ? c : MODE == 1 ? (nmda ? c + boost : c) : (nmda ? 1u : 0u);
}
// K items of a half's list (K = 8 or 4): their W3 words (at byte offset
// woff within the resource, plus the item's line) and spike words are
// requested together; then each branch count is added up and, on the
// branch's last item, thresholded into acc and the NMDA bits bb.
template <uint32_t K, int MODE>
__device__ __forceinline__ void drive_round(const uint32_t* L, __amdgpu_buffer_rsrc_t w3, __amdgpu_buffer_rsrc_t sr,
                                            uint32_t woff, const Params& P, uint32_t& c, uint32_t& acc,
                                            uint32_t& bb) {
    uint32_t it[K];
    const uint4 i0 = *reinterpret_cast<const uint4*>(L);
    it[0] = i0.x; it[1] = i0.y; it[2] = i0.z; it[3] = i0.w;
    if constexpr (K == 8) {
        const uint4 i1 = *reinterpret_cast<const uint4*>(L + 4u);
        it[4] = i1.x; it[5] = i1.y; it[6] = i1.z; it[7] = i1.w;
    }
    uint32_t wv[K], sv[K];
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
        wv[k] = buf_ld(w3, ((it[k] & 0x3FFu) << 7) + woff);
        sv[k] = buf_ld(sr, ((it[k] >> 11) & 0xFFFFFu) << 2);
    }
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
        const uint32_t e = it[k];
        c += popc(wv[k] & sv[k]);
        if ((int32_t)e < 0) {  // last item of the branch: its count is complete
            const bool nmda = c >= P.dthr;
            if (nmda) bb |= 1u << ((e & 0xFFu) >> 3);
            acc += branch_out<MODE>(c, nmda, P.boost);
            c = 0u;
        }
    }
}

// Propagation (k_drive, mm:150-190) for one column per wave, the form for a
// single network, where a timestep is latency-bound: the column's active
// branches are dealt alternately to the two half-waves, halving the items a
// half works through.
template <int MODE>
__device__ __forceinline__ void drive_col(const NetDev& d, const Params& P, const uint32_t* __restrict__ spkIn,
                                          uint32_t col, const DriveIn& in, const uint8_t* fired,
                                          uint32_t* items) {
    const uint32_t lane = lane_id();
    const bool hi = lane >= 32u;
    const uint32_t n = lane & 31u;
    const uint32_t base = col * NSLOT;
    const uint32_t gid = base + n;
    const uint32_t gl = in.gl;
    const uint32_t biasw = (P.useBias != 0u) ? in.biasw : 0u;
    const uint32_t prev = (P.facil != 0u) ? in.prev : 0u;

    // lane s < 32: which of slot s's eight presynaptic columns spiked
    uint32_t fwl;
    if (fired) {
        fwl = byte_mask(*reinterpret_cast<const uint64_t*>(fired + 8u * gl));
    } else {
        const uint4 a = reinterpret_cast<const uint4*>(spkIn)[2u * gl];
        const uint4 b = reinterpret_cast<const uint4*>(spkIn)[2u * gl + 1u];
        fwl = (a.x ? 1u : 0u) | (a.y ? 2u : 0u) | (a.z ? 4u : 0u) | (a.w ? 8u : 0u) | (b.x ? 16u : 0u) |
              (b.y ? 32u : 0u) | (b.z ? 64u : 0u) | (b.w ? 128u : 0u);
    }
    const uint32_t mask = ballot32(fwl != 0u);
    if (!hi) d.firedW[base + n] = (uint8_t)fwl;
    const uint32_t nActive = popc(mask);

    // The active branches are dealt alternately to the two half-waves; each
    // half lists all items of its branches, in order, in LDS (lane s writes
    // branch s's items at a scanned offset), so a single dword load
    // instruction serves one item of each half and a branch's count c
    // (summed over its items) completes inside one half.
    // Branches go to the halves in lane order, split where the column's items
    // halve (both halves compute the same scan).
    const uint32_t ni = popc(fwl);
    const uint32_t sc = half_scan(ni);
    const uint32_t total = __builtin_amdgcn_readlane(sc, 31);
    const uint32_t ex = sc - ni;
    const uint32_t hh = (2u * ex + ni > total) ? 1u : 0u;  // non-decreasing in lane order
    const uint32_t b1 = ballot32(hh != 0u);
    const uint32_t nA = b1 != 0u ? __builtin_amdgcn_readlane(ex, __builtin_ctz(b1)) : total;
    const uint32_t nB = total - nA;
    if (!hi && fwl != 0u) {
        uint32_t* L = items + hh * ITEMS_PER_HALF;
        uint32_t pos = hh ? ex - nA : ex;
        uint32_t m = fwl;
        while (m) {
            const uint32_t w = __builtin_ctz(m);
            m &= m - 1u;
            L[pos++] = drive_item(n, w, gl, m == 0u);
        }
    }
    constexpr uint32_t DB = 8;  // items per half fetched per round trip
    const uint32_t nLines = nA + nB;
    const uint32_t myN = hi ? nB : nA;
    uint32_t* L = items + (hi ? ITEMS_PER_HALF : 0u);  // read after this wave's writes above
    const uint32_t R = (max(nA, nB) + 3u) & ~3u;  // rounds of DB items, then one of 4 if needed
    for (uint32_t i = myN + n; i < R; i += 32u) L[i] = DRIVE_PAD_ITEM;
    const uint32_t ubase = __builtin_amdgcn_readfirstlane(base);  // wave-uniform: the resources live in SGPRs
    const __amdgpu_buffer_rsrc_t w3 = buf_rsrc(d.W3 + (size_t)ubase * TILE_WORDS, NSLOT * TILE_WORDS * 4u);
    const __amdgpu_buffer_rsrc_t sr = buf_rsrc(spkIn, d.ncol * 4u);
    uint32_t acc = 0u, bb = 0u, c = 0u;
    uint32_t r = 0;
    for (; r + DB <= R; r += DB) drive_round<DB, MODE>(L + r, w3, sr, n * 4u, P, c, acc, bb);
    if (r < R) drive_round<4, MODE>(L + r, w3, sr, n * 4u, P, c, acc, bb);
    acc += __shfl_xor(acc, 32);  // the two halves' branches
    bb |= __shfl_xor(bb, 32);
    if (P.facil != 0u && ((prev >> n) & 1u) != 0u) acc += P.facil;
    const uint32_t bi = biasw >> P.biasShift;
    acc = (acc > bi) ? (acc - bi) : 0u;
    if (!hi) {
        if (d.acc16) reinterpret_cast<uint16_t*>(d.acc)[gid] = (uint16_t)acc;
        else d.acc[gid] = acc;
        d.bbits[gid] = bb;
    }
    const uint32_t drv = half_scan(acc);
    if (lane == 31u) {
        d.drive[col] = drv;
        reinterpret_cast<uint2*>(d.cntProp)[col] = make_uint2(in.cnt.x + nActive, in.cnt.y + nLines);
    }
}

constexpr uint32_t ITEMS_PER_COL = NSLOT * 8u;  // LDS item list capacity of a column

// Propagation (k_drive, mm:150-190) for two columns per wave, the form for
// many networks at once, where throughput counts: half h of the wave is
// column colA + h, and lane n of a half is neuron n (and, for the per-branch
// lookups, branch n); no work is duplicated between the halves.  Each half
// lists its column's items in branch order in LDS (lane n writes branch n's
// items at a scanned offset); one dword load instruction then serves one
// item of each half, and a branch's count c, summed over its items,
// completes inside the half.
template <int MODE>
__device__ __forceinline__ void drive_pair(const NetDev& d, const Params& P, const uint32_t* __restrict__ spkIn,
                                           uint32_t colA, const DriveIn& in, const uint8_t* fired,
                                           uint32_t* items) {
    const uint32_t lane = lane_id();
    const uint32_t h = lane >> 5, n = lane & 31u;
    const bool hi = h != 0u;
    const uint32_t col = colA + h;
    const uint32_t gid = col * NPC + n;
    const uint32_t gl = in.gl;
    const uint32_t biasw = (P.useBias != 0u) ? in.biasw : 0u;
    const uint32_t prev = (P.facil != 0u) ? in.prev : 0u;

    // which of branch n's eight presynaptic columns spiked
    uint32_t fwl;
    if (fired) {
        fwl = byte_mask(*reinterpret_cast<const uint64_t*>(fired + 8u * gl));
    } else {
        const uint4 a = reinterpret_cast<const uint4*>(spkIn)[2u * gl];
        const uint4 b = reinterpret_cast<const uint4*>(spkIn)[2u * gl + 1u];
        fwl = (a.x ? 1u : 0u) | (a.y ? 2u : 0u) | (a.z ? 4u : 0u) | (a.w ? 8u : 0u) | (b.x ? 16u : 0u) |
              (b.y ? 32u : 0u) | (b.z ? 64u : 0u) | (b.w ? 128u : 0u);
    }
    d.firedW[col * NSLOT + n] = (uint8_t)fwl;
    const uint32_t mask = (uint32_t)(__ballot(fwl != 0u) >> (32u * h));  // this column's active branches
    const uint32_t nActive = popc(mask);

    const uint32_t ni = popc(fwl);
    const uint32_t sc = half_scan(ni);  // inclusive scan of the item counts over the half's branches
    const uint32_t nA = __builtin_amdgcn_readlane(sc, 31), nB = __builtin_amdgcn_readlane(sc, 63);
    uint32_t* L = items + h * ITEMS_PER_COL;  // written here, read back below by the same wave
    if (fwl != 0u) {
        uint32_t pos = sc - ni;
        uint32_t m = fwl;
        while (m) {
            const uint32_t w = __builtin_ctz(m);
            m &= m - 1u;
            L[pos++] = drive_item(n, w, gl, m == 0u);
        }
    }
    constexpr uint32_t DB = 8;  // items per half fetched per round trip
    const uint32_t myN = hi ? nB : nA;
    const uint32_t R = (max(nA, nB) + 3u) & ~3u;  // rounds of DB items, then one of 4 if needed
    for (uint32_t i = myN + n; i < R; i += 32u) L[i] = DRIVE_PAD_ITEM;
    // the two columns' 64 tiles of W3 (64 KiB) and the spike words, as range-checked resources
    const uint32_t ubase = __builtin_amdgcn_readfirstlane(colA * NSLOT);
    const __amdgpu_buffer_rsrc_t w3 = buf_rsrc(d.W3 + (size_t)ubase * TILE_WORDS, 2u * NSLOT * TILE_WORDS * 4u);
    const __amdgpu_buffer_rsrc_t sr = buf_rsrc(spkIn, d.ncol * 4u);
    const uint32_t n4h = h * (NSLOT * TILE_WORDS * 4u) + n * 4u;
    uint32_t acc = 0u, bb = 0u, c = 0u;
    uint32_t r = 0;
    for (; r + DB <= R; r += DB) drive_round<DB, MODE>(L + r, w3, sr, n4h, P, c, acc, bb);
    if (r < R) drive_round<4, MODE>(L + r, w3, sr, n4h, P, c, acc, bb);
    if (P.facil != 0u && ((prev >> n) & 1u) != 0u) acc += P.facil;
    const uint32_t bi = biasw >> P.biasShift;
    acc = (acc > bi) ? (acc - bi) : 0u;
    if (d.acc16) reinterpret_cast<uint16_t*>(d.acc)[gid] = (uint16_t)acc;
    else d.acc[gid] = acc;
    d.bbits[gid] = bb;
    const uint32_t drv = half_scan(acc);
    if (n == 31u) {
        d.drive[col] = drv;
        // only this column's half writes its counters: a plain read-modify-write
        reinterpret_cast<uint2*>(d.cntProp)[col] = make_uint2(in.cnt.x + nActive, in.cnt.y + myN);
    }
}

constexpr uint32_t FIRED_LDS_COLS = 16384;  // up to this many columns the fired bytes are staged in LDS
// Dynamic LDS bytes of the fired-byte stage (0: read from global memory).
inline uint32_t fired_lds_bytes(uint32_t ncol) { return ncol <= FIRED_LDS_COLS ? (ncol + 15u) / 16u * 16u : 0u; }

// Stage the fired bytes of all columns in LDS while this wave's conn entries
// load.  sync: end with a workgroup barrier even when nothing is staged.
// stim0 (not null): the previous step is a step 0 clamped to the stimulus
// stim0, which lies inside the first inCols0 columns: its spiking-column
// bytes are built from those columns' stimulus words instead.
__device__ __forceinline__ const uint8_t* stage_fired(const NetDev& d, uint4* lds, bool sync = false,
                                                     const uint32_t* stim0 = nullptr, uint32_t inCols0 = 0u) {
    if (d.ncol > FIRED_LDS_COLS) {  // uniform over the grid
        if (sync) __syncthreads();
        return nullptr;
    }
    const uint32_t rows = (d.ncol + 15u) / 16u;
    if (!stim0) {
        const uint4* src = reinterpret_cast<const uint4*>(d.colNZ);
        for (uint32_t i = threadIdx.x; i < rows; i += blockDim.x) lds[i] = src[i];
    } else {
        const uint4* sw = reinterpret_cast<const uint4*>(stim0);  // 4 columns per load
        for (uint32_t i = threadIdx.x; i < rows; i += blockDim.x) {
            uint32_t w[4] = {0u, 0u, 0u, 0u};
            if (16u * i < inCols0) {  // inCols0 is a multiple of 8; a row past it holds no stimulus
#pragma unroll
                for (uint32_t k = 0; k < 4u; ++k) {
                    const uint4 v = 16u * i + 4u * k < inCols0 ? sw[4u * i + k] : make_uint4(0u, 0u, 0u, 0u);
                    w[k] = (v.x ? 1u : 0u) | (v.y ? 0x100u : 0u) | (v.z ? 0x10000u : 0u) | (v.w ? 0x1000000u : 0u);
                }
            }
            lds[i] = make_uint4(w[0], w[1], w[2], w[3]);
        }
    }
    __syncthreads();
    return reinterpret_cast<const uint8_t*>(lds);
}

// ============================================================== column top-K
// theta = target-th largest drive over recurrent columns by three byte-wise
// histogram passes (k_topk_par mm:301-352, k_topk_serial mm:259-292).  One
// workgroup of NT threads per instance.  Up to 4096 values are loaded once
// into registers (4096 / NT per thread) and kept for every pass; further
// values are re-read per pass.  Histogram bins are direct LDS atomics into two
// alternating buffers (each pass clears the buffer the next pass fills, so no
// pass needs a barrier of its own for clearing); the 256-bin suffix scan and
// the crossing bin are computed by wave 0 in registers.
// Returns max(theta, 1) to every thread of the block; thread 0 also stores
// (theta, need) to `theta` when it is not null.
template <uint32_t NT>
__device__ uint32_t topk_block(const uint32_t* __restrict__ drive, uint32_t* __restrict__ theta, uint32_t ncol,
                               uint32_t inCols, uint32_t target, uint32_t par) {
    constexpr uint32_t VPT = 4096u / NT;
    __shared__ uint32_t hist[2][256];
    __shared__ uint32_t sh_need, sh_prefix, sh_or[NT / 64u];
    const uint32_t tid = threadIdx.x, lane = lane_id();
    uint32_t rv[VPT];
#pragma unroll
    for (uint32_t i = 0; i < VPT; ++i) {
        const uint32_t c = inCols + tid + i * NT;
        rv[i] = c < ncol ? drive[c] : 0u;
    }
    const uint32_t cRest = inCols + VPT * NT;  // values beyond the register cache
    for (uint32_t i = tid; i < 256u; i += NT) hist[0][i] = 0u;
    if (tid == 0) { sh_need = target; sh_prefix = 0u; }
    // A leading pass whose byte is zero for every value puts them all in bin 0
    // and leaves (prefix, need) unchanged under both selection rules (except
    // the serial rule with need == 0, which picks bin 255), so it is skipped.
    uint32_t orv = 0;
#pragma unroll
    for (uint32_t i = 0; i < VPT; ++i) orv |= rv[i];
    for (uint32_t c = cRest + tid; c < ncol; c += NT) orv |= drive[c];
    orv = __reduce_or_sync(~0ull, orv);
    if (lane == 0u) sh_or[tid >> 6] = orv;
    __syncthreads();
    uint32_t allOr = 0u;
#pragma unroll
    for (uint32_t w = 0; w < NT / 64u; ++w) allOr |= sh_or[w];
    int byStart = 2;
    if (par || target > 0u) byStart = (allOr >> 16) ? 2 : (allOr >> 8) ? 1 : 0;
    uint32_t hb = 0;
    for (int by = byStart; by >= 0; --by, hb ^= 1u) {
        const uint32_t need = sh_need, prefix = sh_prefix;
        const uint32_t shHi = (uint32_t)(by + 1) * 8u, shLo = (uint32_t)by * 8u;
        uint32_t* __restrict__ h = hist[hb];
        // the other buffer was last read before the previous barrier: clear it for the next pass
        for (uint32_t i = tid; i < 256u; i += NT) hist[hb ^ 1u][i] = 0u;
#pragma unroll
        for (uint32_t i = 0; i < VPT; ++i) {
            const uint32_t c = inCols + tid + i * NT;
            if (c < ncol && (rv[i] >> shHi) == (prefix >> shHi)) atomicAdd(&h[(rv[i] >> shLo) & 255u], 1u);
        }
        for (uint32_t c = cRest + tid; c < ncol; c += NT) {
            const uint32_t dv = drive[c];
            if ((dv >> shHi) == (prefix >> shHi)) atomicAdd(&h[(dv >> shLo) & 255u], 1u);
        }
        __syncthreads();
        if (tid < 64u) {
            const uint32_t h0 = h[4 * lane], h1 = h[4 * lane + 1], h2 = h[4 * lane + 2], h3 = h[4 * lane + 3];
            const uint32_t loc = h0 + h1 + h2 + h3;
            const uint32_t P = wave_scan(loc);  // inclusive prefix over lanes
            const uint32_t S = __builtin_amdgcn_readlane(P, 63) - P + loc;  // inclusive suffix sum over lanes
            const uint32_t above4 = S - loc;
            const uint32_t sc3 = h3 + above4, sc2 = h2 + sc3, sc1 = h1 + sc2, sc0 = S;
            // largest bin i with scan[i] >= need: the suffix sums do not increase with i,
            // so it lies in the last lane whose first bin qualifies
            int best = -1;
            const uint64_t ge = __ballot(sc0 >= need);
            if (ge != 0ull) {
                const uint32_t L = 63u - (uint32_t)__builtin_clzll((unsigned long long)ge);
                best = (int)(4u * L) + (__builtin_amdgcn_readlane(sc1, L) >= need ? 1 : 0) +
                       (__builtin_amdgcn_readlane(sc2, L) >= need ? 1 : 0) +
                       (__builtin_amdgcn_readlane(sc3, L) >= need ? 1 : 0);
            }
            auto scan_at = [&](uint32_t i) -> uint32_t {  // i < 256, wave-uniform
                const uint32_t q = i & 3u, src = i >> 2;
                const uint32_t v = q == 0 ? sc0 : q == 1 ? sc1 : q == 2 ? sc2 : sc3;
                return __builtin_amdgcn_readlane(v, src);
            };
            uint32_t b, above;
            if (par) {
                if (best >= 0 && need > 0u) {
                    b = (uint32_t)best;
                    above = b + 1u < 256u ? scan_at(b + 1u) : 0u;
                } else {
                    b = 0u;
                    above = scan_at(1u);
                }
            } else {
                if (best >= 0) {
                    b = (uint32_t)best;
                    above = b + 1u < 256u ? scan_at(b + 1u) : 0u;
                } else {
                    b = 0u;
                    above = 0u;
                }
            }
            if (tid == 0) { sh_need = need - above; sh_prefix = prefix | (b << shLo); }
        }
        __syncthreads();
    }
    const uint32_t th = max(sh_prefix, 1u);
    if (tid == 0 && theta) { theta[0] = th; theta[1] = sh_need; }
    return th;
}

__global__ __launch_bounds__(1024) void k_topk_test(const uint32_t* __restrict__ drive, uint32_t* theta,
                                                   uint32_t ncol, uint32_t inCols, uint32_t target, uint32_t par) {
    topk_block<1024>(drive + (size_t)blockIdx.x * ncol, theta + 2 * blockIdx.x, ncol, inCols, target, par);
}

// ============================================================== readout + plasticity
// k_emit (mm:192-236) and k_bias (mm:240-253) fused: everything after the
// global theta is local to a column.  A workgroup owns 16 columns, one wave
// each.  The k-of-32 selection is computed only where a spike can leave (gate
// passed or input clamped); every other column is silent, and its margin
// flag -- which plasticity reads only for spiking columns -- is written as 0.
// The wave also records whether its column fired (read by the next
// propagation).
// STATS = false drops the side outputs (stats and drive-log atomics, LTP/LTD
// counters); spikes, history, margin, bias and synapses are unaffected.
// Whether a voted tile needs all 8 words rewritten: the update touches words
// of silent presynaptic columns (the legacy rule with depression, hetero-
// synaptic LTD), or the band clip may still move counters the tile's spikes
// do not touch.  Otherwise only the spiking columns' words can change.
__device__ __forceinline__ bool full_tile_update(const Params& P, uint32_t band) {
    return P.phase == 0u || (P.phase == 1u && P.hetero != 0u) || (P.bandHi != 0u && band == 0u);
}
// Whether a column learns at a learning timestep of parameters P (mm:440-455):
// plasticity is on in this phase, the column spiked (spike word b) and, with
// margin gating, its k-of-32 margin flag is set.
__device__ __forceinline__ bool learns(const Params& P, uint32_t b, uint32_t marg) {
    return P.phase != 3u && b != 0u && !(P.marginGate != 0u && marg == 0u);
}

// With side statistics a learning column is listed for k_learn, which
// rewrites all of the column's voted tiles in one workgroup (the per-neuron
// LTP/LTD totals need them together); its work counters are added here.
__device__ __forceinline__ void list_learning_column(const NetDev& d, const Params& P, uint32_t col,
                                                     uint32_t bbv, uint32_t fwb, uint32_t band, uint2 lc) {
    const uint32_t lane = lane_id();
    const uint32_t par = P.step & 1u;
    const uint32_t voted = half_or(bbv);
    const bool v = lane < 32u && ((voted >> lane) & 1u) != 0u;
    const uint32_t lines = wave_sum(v ? (full_tile_update(P, band) ? 8u : popc(fwb)) : 0u);
    if (lane == 0u) {
        reinterpret_cast<uint2*>(d.cntLearn)[col] = make_uint2(lc.x + popc(voted), lc.y + lines);
        const uint32_t l = list_of(col), cap = list_cols(d.ncol);
        d.learnList[(par * NLISTS + l) * cap + atomicAdd(list_count(d, par, l), 1u)] = col;
    }
}

// A column's emit inputs.  Every column needs its drive, accumulators, clamp
// word and bias (read before the column gate, which they overlap).  With side
// statistics a column that can spike (gate passed or input clamped) also
// needs its NMDA bits and, when learning, its plasticity inputs (fired
// masks, band flags, counters): those are requested right after the gate, so
// their round trip overlaps the k-of-32 selection, and the other columns
// never read them.  Without side statistics plasticity reads its inputs
// itself (vote_column), one timestep later.
struct EmitIn { uint32_t dv, v, stimW, bbv, fwb, band, bia; uint2 lc; };
template <bool LEARN, bool STATS>
__device__ __forceinline__ EmitIn emit_inputs(const NetDev& d, const Params& P, uint32_t col) {
    const uint32_t n = lane_id() & 31u;
    const uint32_t gid = col * NPC + n;
    const bool clampCol = P.clamp != 0u && col < P.inCols;
    const bool rec = col >= P.inCols;
    EmitIn in;
    in.dv = d.drive[col];
    in.v = d.acc16 ? (uint32_t)reinterpret_cast<const uint16_t*>(d.acc)[gid] : d.acc[gid];
    in.stimW = clampCol ? d.stim[col] : 0u;
    in.bbv = 0u; in.fwb = 0u; in.band = 0u;
    in.lc = make_uint2(0u, 0u);
    in.bia = (LEARN && rec) ? d.bias[gid] : 0u;
    return in;
}
template <bool LEARN, bool STATS>
__device__ __forceinline__ void emit_inputs_spiking(const NetDev& d, const Params& P, uint32_t col, EmitIn& in) {
    if (!STATS) return;
    const uint32_t n = lane_id() & 31u;
    in.bbv = d.bbits[col * NPC + n];
    if (LEARN && P.phase != 3u) {
        in.fwb = d.firedW[col * NSLOT + n];
        in.band = d.tileBand[col * NSLOT + n];
        in.lc = reinterpret_cast<const uint2*>(d.cntLearn)[col];
    }
}

// Returns the column's spike word and margin flag (0, 0 past ncol).
template <bool LEARN, bool STATS>
__device__ __forceinline__ uint2 emit_cols(const NetDev& d, const Params& P, const uint32_t* __restrict__ spkIn,
                                           uint32_t* __restrict__ spkOut, uint32_t col, const EmitIn& in,
                                           bool gatePass) {
    const uint32_t lane = lane_id();
    const uint32_t n = lane & 31u;
    const bool hi = lane >= 32u;
    if (col >= d.ncol) return make_uint2(0u, 0u);
    const uint32_t gid = col * NPC + n;
    const bool clampCol = P.clamp != 0u && col < P.inCols;
    const bool rec = col >= P.inCols;
    const uint32_t v = in.v, stimW = in.stimW, bbv = in.bbv, bia = in.bia;
    uint32_t b = 0u, vk = 0u, vk1 = 0u;
    if (clampCol || gatePass) {
        const Kwta kw = kwta32(v, P.k, n);
        vk = kw.vk; vk1 = kw.vk1;
        b = gatePass ? kw.ball : 0u;
        if (clampCol) b = stimW;
    }
    const bool post = ((b >> n) & 1u) != 0u;
    if (STATS) {
        const uint32_t x = (!hi && post && rec) ? popc(bbv) : 0u;
        const bool mem = !hi && rec && ((d.refBall[col] >> n) & 1u) != 0u;
        const bool shared = d.mult[gid] > 1u;
        const uint32_t sumS = wave_sum(mem && shared ? v : 0u);
        const uint32_t sumE = wave_sum(mem && !shared ? v : 0u);
        const uint32_t cntS = popc(ballot32(mem && shared));
        const uint32_t cntE = popc(ballot32(mem && !shared));
        const uint32_t xs = wave_sum(x);
        if (lane == 0u) {
            const uint32_t t = P.step;
            if (xs) atomicAdd(&d.stats[2], xs);
            if (cntS) { atomicAdd(&d.driveLog[t * 4u + 0u], sumS); atomicAdd(&d.driveLog[t * 4u + 1u], cntS); }
            if (cntE) { atomicAdd(&d.driveLog[t * 4u + 2u], sumE); atomicAdd(&d.driveLog[t * 4u + 3u], cntE); }
            if (b != 0u && rec) {
                const uint32_t sb = 64u + t * 8u;
                atomicAdd(&d.stats[0], 1u);
                atomicAdd(&d.stats[1], popc(b));
                atomicAdd(&d.stats[sb + 0u], vk - vk1);
                atomicAdd(&d.stats[sb + 1u], vk);
                atomicAdd(&d.stats[sb + 2u], 1u);
                if (vk == vk1) atomicAdd(&d.stats[sb + 3u], 1u);
                if (vk == 0u) atomicAdd(&d.stats[sb + 4u], 1u);
            }
        }
    }
    if (lane == 0u) {
        spkOut[col] = b;
        d.marg[col] = (vk > vk1) ? 1u : 0u;
        d.colNZ[col] = b != 0u ? 1u : 0u;
        if (P.histOn != 0u) d.hist[(size_t)P.step * d.ncol + col] = b;
    }
    if (LEARN && !hi && rec) {  // k_bias
        const int dec = ((P.step % max(P.biasDownEvery, 1u)) == 0u) ? (int)P.biasDown : 0;
        const int nb = (int)bia + (post ? (int)P.biasUp : -dec);
        d.bias[gid] = (uint32_t)max(0, nb);
    }
    // without side statistics the plasticity of this step runs one timestep later (vote_column), or,
    // at a presentation's last timestep, right after the emit (k_emit_learn)
    if (STATS && LEARN && learns(P, b, (vk > vk1) ? 1u : 0u))
        list_learning_column(d, P, col, post ? bbv : 0u, in.fwb, in.band, in.lc);
    return make_uint2(b, (vk > vk1) ? 1u : 0u);
}

// One voted tile of k_learn (mm:436-495).  Lane (h, n) owns words 4h..4h+3
// of neuron n in each of the four planes; with `fl` bit 8 clear only the
// words of presynaptic columns that spiked (fl bits 0-7) are read and
// written, the others being unchanged by the update.  A full pass that
// applies the band clip leaves every counter of the tile inside the band.
__device__ __forceinline__ void learn_tile(const NetDev& d, const Params& P, const uint32_t* __restrict__ spkIn,
                                           uint32_t tile, uint32_t g, uint32_t wm, uint32_t fl, uint32_t& nLTP,
                                           uint32_t& nLTD) {
    const uint32_t lane = lane_id(), n = lane & 31u;
    const bool hi = lane >= 32u;
    const bool full = (fl & 0x100u) != 0u;
    const uint32_t my = full ? 0xFu : ((fl >> (hi ? 4u : 0u)) & 0xFu);
    const uint4 sp4 = group_words(spkIn, g, hi);
    const uint32_t sp[4] = {sp4.x, sp4.y, sp4.z, sp4.w};
    const size_t t0 = (size_t)__builtin_amdgcn_readfirstlane(tile) * TILE_WORDS;
    const __amdgpu_buffer_rsrc_t r0 = buf_rsrc(d.W0 + t0, TILE_WORDS * 4u), r1 = buf_rsrc(d.W1 + t0, TILE_WORDS * 4u),
                                 r2 = buf_rsrc(d.W2 + t0, TILE_WORDS * 4u), r3 = buf_rsrc(d.W3 + t0, TILE_WORDS * 4u);
    const uint32_t o0 = ((hi ? 4u : 0u) * 32u + n) * 4u;
    uint32_t a[4][4], o[4];
#pragma unroll
    for (uint32_t k = 0; k < 4u; ++k) {  // words this lane leaves alone are neither read nor written
        o[k] = ((my >> k) & 1u) ? o0 + 128u * k : SKIP;
        a[k][0] = buf_ld(r0, o[k]); a[k][1] = buf_ld(r1, o[k]); a[k][2] = buf_ld(r2, o[k]); a[k][3] = buf_ld(r3, o[k]);
    }
    const bool want = ((wm >> n) & 1u) != 0u;
#pragma unroll
    for (uint32_t k = 0; k < 4u; ++k) {
        learn_word(a[k][0], a[k][1], a[k][2], a[k][3], (my >> k) & 1u ? sp[k] : 0u, want, P, nLTP, nLTD);
        buf_st(r0, o[k], a[k][0]); buf_st(r1, o[k], a[k][1]); buf_st(r2, o[k], a[k][2]); buf_st(r3, o[k], a[k][3]);
    }
    if (full && P.bandHi != 0u && lane == 0u) d.tileBand[tile] = 1u;
}

// k_learn for the columns emit listed (side statistics on): one listed column
// per workgroup (grid-stride), its voted tiles dealt round-robin to the four
// waves; the per-neuron LTP/LTD totals are reduced in LDS (mm:491-494).
template <bool STATS>
__device__ __forceinline__ void learn_cols(const NetDev& d, const Params& P, const uint32_t* __restrict__ spkIn,
                                           const uint32_t* __restrict__ spkOut) {
    __shared__ uint32_t sLTP[4][32], sLTD[4][32];
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    const uint32_t n = lane & 31u;
    const bool hi = lane >= 32u;
    const uint32_t par = P.step & 1u;
    const uint32_t l = blockIdx.x % NLISTS, cap = list_cols(d.ncol);  // workgroup b serves list b % NLISTS
    if (blockIdx.x / NLISTS >= gridDim.x / NLISTS) return;  // workgroup-uniform: no barrier is skipped by part of it
    const uint32_t nlist = *list_count(d, par, l);
    for (uint32_t li = blockIdx.x / NLISTS; li < nlist; li += gridDim.x / NLISTS) {
        const uint32_t col = d.learnList[(par * NLISTS + l) * cap + li];
        const uint32_t b = spkOut[col];
        const bool post = ((b >> n) & 1u) != 0u;
        const uint32_t bb = post ? d.bbits[col * NPC + n] : 0u;
        const uint32_t cn = d.conn[col * NSLOT + n];
        const uint32_t fwb = d.firedW[col * NSLOT + n];
        const uint32_t band = d.tileBand[col * NSLOT + n];
        // lane s < 32: the neurons whose slot-s branch learns (the NMDA bits transposed)
        const uint32_t tr = half_transpose32(bb);
        const uint32_t want = lane < 32u ? tr : 0u;
        const uint32_t fl = fwb | (full_tile_update(P, band) ? 0x100u : 0u);
        uint32_t voted = ballot32(want != 0u);
        uint32_t nLTP = 0, nLTD = 0;
        for (uint32_t r = 0; voted; ++r) {
            const uint32_t s = __builtin_ctz(voted);
            voted &= voted - 1u;
            if ((r & 3u) != wave) continue;
            learn_tile(d, P, spkIn, col * NSLOT + s, __builtin_amdgcn_readlane(cn, s),
                       __builtin_amdgcn_readlane(want, s), __builtin_amdgcn_readlane(fl, s), nLTP, nLTD);
        }
        if (STATS) {  // per-neuron totals over both halves and all four waves (mm:491-494)
            nLTP += __shfl_xor(nLTP, 32);
            nLTD += __shfl_xor(nLTD, 32);
            if (!hi) { sLTP[wave][n] = nLTP; sLTD[wave][n] = nLTD; }
            __syncthreads();
            if (wave == 0u) {
                const uint32_t tp = hi ? 0u : sLTP[0][n] + sLTP[1][n] + sLTP[2][n] + sLTP[3][n];
                const uint32_t td = hi ? 0u : sLTD[0][n] + sLTD[1][n] + sLTD[2][n] + sLTD[3][n];
                const uint32_t sp_ = wave_sum(tp), sd_ = wave_sum(td);
                const uint32_t cp = popc(ballot32(tp != 0u)), cd = popc(ballot32(td != 0u));
                if (lane == 0u) {
                    if (sp_) atomicAdd(&d.stats[8], sp_);
                    if (sd_) atomicAdd(&d.stats[9], sd_);
                    if (cp) atomicAdd(&d.stats[10], cp);
                    if (cd) atomicAdd(&d.stats[11], cd);
                }
            }
            __syncthreads();
        }
    }
}

// ---- plasticity without side statistics, one timestep late ----
// The plasticity of timestep t (k_learn, mm:436-495) rewrites tiles (column,
// branch) of learning columns only, and the next reader of a tile's words is
// its own column's propagation at timestep t + 1.  So the rewrite runs at the
// start of that propagation, in the workgroup that propagates the column
// (spread over the workgroup's waves, then a workgroup barrier before the
// column's lines are read), and needs no launch of its own.  Everything the
// update reads is still in place then: the column's spike word and margin
// flag (step t's emit), its NMDA bits and fired-column masks (step t's
// propagation; step t + 1's overwrites them only after this), the band flags,
// and step t's presynaptic spikes (the ping-pong buffer that step t + 1's emit
// overwrites).  The last timestep of a presentation has no propagation after
// it: its plasticity runs in its emit launch instead (k_emit_learn).
template <uint32_t NC>  // columns of the workgroup
struct VoteList {
    uint4 tile[NC * NSLOT];         // voted tiles: {tile, presynaptic group, learning-neuron mask, 0}
    uint16_t line[NC * NSLOT * 8];  // word lines to rewrite: (tile entry << 3) | word
    uint32_t nt, nl;                // entries listed
};

// A column's plasticity inputs, requested with the launch's first loads (for
// every column: they are a few hundred bytes each, and waiting for the spike
// word first would cost a round trip): NMDA bits, fired-column masks and band
// flags of lane n's neuron / branch slot, the margin flag and the column's
// plasticity work counters `cnt` (tiles, lines).
struct VoteIn { uint32_t bb, fwb, band, marg; uint2 lc; };
__device__ __forceinline__ VoteIn vote_inputs(const NetDev& d, uint32_t col, const uint2* __restrict__ cnt) {
    const uint32_t n = lane_id() & 31u;
    return VoteIn{d.bbits[col * NPC + n], d.firedW[col * NSLOT + n], d.tileBand[col * NSLOT + n], d.marg[col],
                  cnt[col]};
}

// Lists the voted tiles of a learning column, and their word lines, into `vl`
// and adds them to the column's plasticity work counters `cnt`.  Branch slot
// s is voted (mm:446-452) when some spiking neuron has its slot-s NMDA bit
// set; a voted tile rewrites all 8 word lines when the update can touch words
// of silent presynaptic columns (full_tile_update), else only the lines of
// the presynaptic columns that spiked.  A full pass with the band clip
// leaves every counter of the tile inside the band, which its band flag
// records (read at the next voting).  Run by the half-wave h that holds the
// column: lane n of the half is neuron n and branch slot n, gl =
// conn[col * 32 + n], b the column's spike word at the learning timestep.
template <uint32_t NC>
__device__ __forceinline__ void vote_column(const NetDev& d, const Params& PL, uint32_t col, uint32_t b,
                                            uint32_t gl, const VoteIn& vi, uint32_t h, uint2* __restrict__ cnt,
                                            VoteList<NC>& vl) {
    const uint32_t n = lane_id() & 31u;
    const uint32_t want = half_transpose32(((b >> n) & 1u) != 0u ? vi.bb : 0u);  // lane s: neurons whose slot-s branch learns
    const uint32_t voted = hball(want != 0u, h);
    const bool full = full_tile_update(PL, vi.band);
    const uint32_t lm = want != 0u ? (full ? 0xFFu : (vi.fwb & 0xFFu)) : 0u;  // the tile's lines to rewrite
    const uint32_t nl = popc(lm);
    const uint32_t lsc = half_scan(nl);
    const uint32_t ltot = __shfl(lsc, (int)(32u * h + 31u));
    uint2 base = make_uint2(0u, 0u);
    if (n == 0u) {
        base = make_uint2(atomicAdd(&vl.nt, popc(voted)), atomicAdd(&vl.nl, ltot));
        cnt[col] = make_uint2(vi.lc.x + popc(voted), vi.lc.y + ltot);  // only this half writes them
    }
    base.x = __shfl(base.x, (int)(32u * h));
    base.y = __shfl(base.y, (int)(32u * h));
    if (want != 0u) {
        const uint32_t i = base.x + popc(voted & ((1u << n) - 1u));
        vl.tile[i] = make_uint4(col * NSLOT + n, gl, want, 0u);
        uint32_t pos = base.y + lsc - nl;
        for (uint32_t m = lm; m; m &= m - 1u) vl.line[pos++] = (uint16_t)((i << 3) | (uint32_t)__builtin_ctz(m));
        if (full && PL.bandHi != 0u) d.tileBand[col * NSLOT + n] = 1u;
    }
}

// One round of a half-wave's lines (slot u: line j0 + q + u * nq of the list).
template <uint32_t U>
struct LineRound { uint32_t off[U], wm[U], sp[U], a[U][4]; };
struct PlaneRsrc { __amdgpu_buffer_rsrc_t r0, r1, r2, r3, rs; };

// Slot u of the wave holds a line for its first half (and so for any of its
// lanes) while j0 + 2 * wave + u * nq < L: wave-uniform, so the slots past
// the list cost no instructions.  `left` = L - j0 - 2 * wave, which wraps
// past L when the wave has no slot left.
__device__ __forceinline__ bool slot_live(uint32_t u, uint32_t nq, uint32_t left, uint32_t L) {
    return u * nq < left && left <= L;
}

// Requests round j0's plane words and spike words (no request for a slot
// past the list: its offsets are SKIP, it reads 0 and writes nothing).
template <uint32_t NC, uint32_t U>
__device__ __forceinline__ void request_lines(LineRound<U>& R, const VoteList<NC>& vl, const PlaneRsrc& pr,
                                              uint32_t j0, uint32_t L, uint32_t q, uint32_t nq, uint32_t wave,
                                              uint32_t tile0) {
    const uint32_t n = lane_id() & 31u;
    const uint32_t left = L - j0 - 2u * wave;
    uint32_t e[U], so[U];
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
        if (!slot_live(u, nq, left, L)) break;
        e[u] = vl.line[min(j0 + q + u * nq, L - 1u)];
    }
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
        if (!slot_live(u, nq, left, L)) break;
        const uint4 te = vl.tile[e[u] >> 3];
        const uint32_t w = e[u] & 7u;
        const bool ok = j0 + q + u * nq < L;
        R.off[u] = ok ? (((te.x - tile0) * 8u + w) * 32u + n) * 4u : SKIP;
        so[u] = ok ? (8u * te.y + w) * 4u : SKIP;
        R.wm[u] = te.z;
    }
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
        if (!slot_live(u, nq, left, L)) break;
        R.a[u][0] = buf_ld(pr.r0, R.off[u]); R.a[u][1] = buf_ld(pr.r1, R.off[u]);
        R.a[u][2] = buf_ld(pr.r2, R.off[u]); R.a[u][3] = buf_ld(pr.r3, R.off[u]);
        R.sp[u] = buf_ld(pr.rs, so[u]);
    }
}

// Applies k_learn's per-word update to round j0's lines and writes them back.
template <uint32_t U>
__device__ __forceinline__ void update_lines(LineRound<U>& R, const PlaneRsrc& pr, const Params& PL, uint32_t j0,
                                             uint32_t L, uint32_t nq, uint32_t wave) {
    const uint32_t n = lane_id() & 31u;
    const uint32_t left = L - j0 - 2u * wave;
    uint32_t nLTP = 0u, nLTD = 0u;  // not reported without side statistics
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
        if (!slot_live(u, nq, left, L)) break;
        learn_word(R.a[u][0], R.a[u][1], R.a[u][2], R.a[u][3], R.sp[u], ((R.wm[u] >> n) & 1u) != 0u, PL, nLTP, nLTD);
        buf_st(pr.r0, R.off[u], R.a[u][0]); buf_st(pr.r1, R.off[u], R.a[u][1]);
        buf_st(pr.r2, R.off[u], R.a[u][2]); buf_st(pr.r3, R.off[u], R.a[u][3]);
    }
}

// The ROUNDS_EAGER form of rewrite_voted (below): one round at a time, every
// slot's list entries read up front, words requested for the live slots only.
template <uint32_t NC, uint32_t U, uint32_t SPAN>
__device__ __forceinline__ void rewrite_rounds_eager(const NetDev& d, const Params& PL, const uint32_t* __restrict__ spkL,
                                              const VoteList<NC>& vl, uint32_t wave, uint32_t nwaves,
                                              uint32_t tile0) {
    const uint32_t lane = lane_id(), n = lane & 31u;
    const uint32_t nq = 2u * nwaves, q = 2u * wave + (lane >> 5);
    const uint32_t L = vl.nl;
    const size_t t0 = (size_t)__builtin_amdgcn_readfirstlane(tile0) * TILE_WORDS;
    constexpr uint32_t BYTES = SPAN * NSLOT * TILE_WORDS * 4u;
    const __amdgpu_buffer_rsrc_t r0 = buf_rsrc(d.W0 + t0, BYTES), r1 = buf_rsrc(d.W1 + t0, BYTES),
                                 r2 = buf_rsrc(d.W2 + t0, BYTES), r3 = buf_rsrc(d.W3 + t0, BYTES);
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(spkL, d.ncol * 4u);
    uint32_t nLTP = 0u, nLTD = 0u;  // not reported without side statistics
    for (uint32_t j0 = 0; j0 < L; j0 += nq * U) {  // L is workgroup-uniform
        uint32_t e[U], off[U], so[U], wm[U], a[U][4], sp[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) e[u] = vl.line[min(j0 + q + u * nq, L - 1u)];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const uint4 te = vl.tile[e[u] >> 3];
            const uint32_t w = e[u] & 7u;
            const bool ok = j0 + q + u * nq < L;  // else no request: reads 0, writes nothing
            off[u] = ok ? (((te.x - tile0) * 8u + w) * 32u + n) * 4u : SKIP;
            so[u] = ok ? (8u * te.y + w) * 4u : SKIP;
            wm[u] = te.z;
        }
        // slot u of this wave holds a line for its first half (and so for any of its lanes) while
        // j0 + 2 * wave + u * nq < L: wave-uniform, so the slots past the list cost no instructions
        const uint32_t left = L - j0 - 2u * wave;  // wraps when this wave has no slot left
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            if (u * nq >= left || left > L) break;
            a[u][0] = buf_ld(r0, off[u]); a[u][1] = buf_ld(r1, off[u]);
            a[u][2] = buf_ld(r2, off[u]); a[u][3] = buf_ld(r3, off[u]);
            sp[u] = buf_ld(rs, so[u]);
        }
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            if (u * nq >= left || left > L) break;
            learn_word(a[u][0], a[u][1], a[u][2], a[u][3], sp[u], ((wm[u] >> n) & 1u) != 0u, PL, nLTP, nLTD);
            buf_st(r0, off[u], a[u][0]); buf_st(r1, off[u], a[u][1]);
            buf_st(r2, off[u], a[u][2]); buf_st(r3, off[u], a[u][3]);
        }
    }
}

// The listed word lines (k_learn's per-word update, mm:471-487), dealt to the
// workgroup's half-waves: half-wave q rewrites lines q, q + nq, q + 2nq, ...,
// U at a time (the U lines' plane words and spike words requested before any
// is updated); lane n of the half is neuron n.  The listed tiles belong to
// the workgroup's NC columns, which lie among the SPAN columns from tile0 on:
// each plane is addressed through a buffer resource over those SPAN * 32
// tiles (base in scalar registers, one 32-bit offset per line for all four
// planes).  FORM (measured per kernel): ROUNDS_EAGER reads every slot's list
// entries up front and requests the live slots' words, one round at a time
// (the batched kernels); ROUNDS one round at a time, live slots only;
// PIPELINED requests the next round's words before this round's are updated
// (two rounds of registers), so a workgroup with many lines waits for memory
// about once instead of once per round (the listed lines are distinct, so no
// round reads a line another writes).
enum RewriteForm : uint32_t { ROUNDS_EAGER, ROUNDS, PIPELINED };
template <uint32_t NC, uint32_t U = 4, uint32_t SPAN = NC, RewriteForm FORM = ROUNDS_EAGER>
__device__ __forceinline__ void rewrite_voted(const NetDev& d, const Params& PL, const uint32_t* __restrict__ spkL,
                                              const VoteList<NC>& vl, uint32_t wave, uint32_t nwaves,
                                              uint32_t tile0) {
    const uint32_t nq = 2u * nwaves, q = 2u * wave + (lane_id() >> 5);
    const uint32_t L = vl.nl;  // workgroup-uniform
    const size_t t0 = (size_t)__builtin_amdgcn_readfirstlane(tile0) * TILE_WORDS;
    constexpr uint32_t BYTES = SPAN * NSLOT * TILE_WORDS * 4u;
    const PlaneRsrc pr{buf_rsrc(d.W0 + t0, BYTES), buf_rsrc(d.W1 + t0, BYTES), buf_rsrc(d.W2 + t0, BYTES),
                       buf_rsrc(d.W3 + t0, BYTES), buf_rsrc(spkL, d.ncol * 4u)};
    const uint32_t R = nq * U;
    if (FORM == ROUNDS) {
        for (uint32_t j0 = 0; j0 < L; j0 += R) {
            LineRound<U> A;
            request_lines<NC, U>(A, vl, pr, j0, L, q, nq, wave, tile0);
            update_lines<U>(A, pr, PL, j0, L, nq, wave);
        }
        return;
    }
    if (FORM == ROUNDS_EAGER) {
        rewrite_rounds_eager<NC, U, SPAN>(d, PL, spkL, vl, wave, nwaves, tile0);
        return;
    }
    LineRound<U> A, B;
    if (L == 0u) return;
    request_lines<NC, U>(A, vl, pr, 0u, L, q, nq, wave, tile0);
    for (uint32_t j0 = 0;; j0 += 2u * R) {
        if (j0 + R < L) request_lines<NC, U>(B, vl, pr, j0 + R, L, q, nq, wave, tile0);
        update_lines<U>(A, pr, PL, j0, L, nq, wave);
        if (j0 + R >= L) break;
        if (j0 + 2u * R < L) request_lines<NC, U>(A, vl, pr, j0 + 2u * R, L, q, nq, wave, tile0);
        update_lines<U>(B, pr, PL, j0 + R, L, nq, wave);
        if (j0 + 2u * R >= L) break;
    }
}

// Propagation of one network, DRIVE_COLS columns per workgroup, one per wave.  FUSE:
// the deferred plasticity of the previous timestep (parameters PL,
// presynaptic spikes spkL) first rewrites the voted tiles of the
// workgroup's columns.  FIRST: this is step 1 of a presentation without side
// statistics whose step 0 (parameters PL) is clamped to a stimulus inside the
// input columns.  Step 0's spikes are then that stimulus (spkIn: d.stim, or a
// row of the pattern table, which this launch also copies into d.stim),
// since with cleared spikes no
// column passes the gate: this launch also does step 0's own work for its
// columns (as k_emit_first: spike word, history row, bias decay), so step 0
// needs no launch.  Step 0's margin and spiking flags are not written: no
// plasticity reads them (step 0 votes no tile) and the next emit rewrites them.
template <int MODE, bool FUSE, bool FIRST = false>
__global__ __launch_bounds__(64 * DRIVE_COLS) void k_drive(NetDev d, Params P, const uint32_t* __restrict__ spkIn,
                                                           Params PL, const uint32_t* __restrict__ spkL) {
    constexpr uint32_t DC = DRIVE_COLS;
    extern __shared__ uint4 fired[];  // fired_lds_bytes(ncol)
    __shared__ __align__(16) uint32_t items[DC][2 * ITEMS_PER_HALF];
    __shared__ VoteList<DC> vl;
    // the learning-column list of step t-1 has been consumed; free it for step t+1
    if (blockIdx.x == 0u && threadIdx.x < NLISTS) *list_count(d, (P.step & 1u) ^ 1u, threadIdx.x) = 0u;
    const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
    const uint32_t col = blockIdx.x * DC + wave;
    const uint32_t ccol = min(col, d.ncol - 1u);
    DriveIn in = drive_inputs(d, spkIn, ccol);  // in.prev: the column's spike word of the previous step
    VoteIn vi{};
    if (FUSE) vi = vote_inputs(d, ccol, reinterpret_cast<const uint2*>(d.cntLearn));
    if (FUSE && threadIdx.x == 0u) { vl.nt = 0u; vl.nl = 0u; }
    if (FIRST) {
        if (PL.learn != 0u && ccol >= PL.inCols) {  // step 0's k_bias: recurrent columns never spike there
            const int dec = ((PL.step % max(PL.biasDownEvery, 1u)) == 0u) ? (int)PL.biasDown : 0;
            in.biasw = (uint32_t)max(0, (int)in.biasw - dec);
            if (col < d.ncol && lane < 32u) d.bias[col * NPC + lane] = in.biasw;
        }
        if (col < d.ncol && lane == 0u) {
            d.spk[1][col] = in.prev;
            if (spkIn != d.stim) d.stim[col] = in.prev;  // a pattern table row: kept for the clamped steps
            if (PL.histOn != 0u) d.hist[(size_t)PL.step * d.ncol + col] = in.prev;
        }
    }
    const uint8_t* f = stage_fired(d, fired, FUSE, FIRST ? spkIn : nullptr, PL.inCols);
    if (FUSE) {
        if (col < d.ncol && lane < 32u && learns(PL, in.prev, vi.marg))
            vote_column<DC>(d, PL, col, in.prev, in.gl, vi, 0u, reinterpret_cast<uint2*>(d.cntLearn), vl);
        __syncthreads();
        if (vl.nl != 0u) {  // workgroup-uniform
            rewrite_voted<DC, 4, DC, PIPELINED>(d, PL, spkL, vl, wave, DC, blockIdx.x * DC * NSLOT);
            __syncthreads();  // the rewritten lines are read below, by any wave of the workgroup
        }
    }
    if (col >= d.ncol) return;
    drive_col<MODE>(d, P, spkIn, col, in, f, items[wave]);
}

// The column gate (drive >= max(theta, 1), mm:211) of a 1024-thread
// workgroup's 16 columns, one per wave, when 1 <= target <= recurrent columns
// <= 4096.  Theta is then the target-th largest recurrent drive under both
// top-K rules (every radix pass has a crossing bin), so a column passes
// exactly when its drive is at least 1 and fewer than `target` recurrent
// drives exceed it.  The workgroup holds the recurrent drives in registers
// (4 per thread); each wave counts, for every one of the 16 columns, the
// drives of its 256 that exceed the column's (ballots), and after one
// barrier each wave adds up its own column's 16 partial counts.  This takes
// the place of the two or three histogram passes of the radix select.
// Column of wave w of emit workgroup b: the workgroups on XCD x = b % 8 take
// the columns that propagation ran on XCD x (its DRIVE_COLS-column
// workgroups j with j % 8 = x, whole or in parts of WG columns), so a
// column's accumulators, NMDA bits and fired masks are read from the L2 that
// received them.  With WG waves
// (columns) per workgroup the grid is ceil(ncol / (8 * WG)) * 8 workgroups;
// waves past ncol emit nothing.
template <uint32_t WG = 16>
__device__ __forceinline__ uint32_t emit_col(uint32_t b, uint32_t w) {
    constexpr uint32_t DC = DRIVE_COLS;
    if (WG >= DC) {  // G whole propagation workgroups per emit workgroup
        constexpr uint32_t G = WG / DC;
        static_assert(WG < DC || G * DC == WG, "whole propagation workgroups");
        const uint32_t j = (b & 7u) + 8u * (G * (b >> 3) + w / DC);
        return DC * j + w % DC;
    }
    constexpr uint32_t H = DC / WG;  // emit workgroups per propagation workgroup
    const uint32_t r = b >> 3, j = (b & 7u) + 8u * (r / H);
    return DC * j + (r % H) * WG + w;
}

// The count behind the gate for the NCW columns of an NT-thread workgroup
// (up to 4096 recurrent drives held in registers): lane j < NCW holds in dl
// the drive of the workgroup's column j; returns, for column `mine`, the
// number of recurrent drives that exceed its drive.  Every wave of the
// workgroup must call it (it ends with a barrier).
template <uint32_t NCW, uint32_t NT = 1024>
__device__ __forceinline__ uint32_t drives_above(const uint32_t* __restrict__ drive, uint32_t ncol, uint32_t inCols,
                                                 uint32_t dl, uint32_t mine) {
    constexpr uint32_t NV = 4096u / NT, NWV = NT / 64u;
    __shared__ uint32_t part[NWV][NCW];  // [wave][column of the workgroup]
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    uint32_t rv[NV];
#pragma unroll
    for (uint32_t i = 0; i < NV; ++i) {
        const uint32_t c = inCols + tid + i * NT;
        rv[i] = c < ncol ? drive[c] : 0u;  // past the end: 0, which exceeds no drive
    }
    uint32_t own = 0u;
#pragma unroll 4
    for (uint32_t j = 0; j < NCW; ++j) {
        const uint32_t dj = __builtin_amdgcn_readlane(dl, j);
        uint32_t cnt = 0u;
#pragma unroll
        for (uint32_t i = 0; i < NV; ++i) cnt += (uint32_t)__popcll(__ballot(rv[i] > dj));
        if (lane == j) own = cnt;
    }
    if (lane < NCW) part[wave][lane] = own;
    __syncthreads();
    uint32_t tot = 0u;
#pragma unroll
    for (uint32_t w = 0; w < NWV; ++w) tot += part[w][mine];
    return tot;
}
template <uint32_t WG = 16>
__device__ __forceinline__ bool gate_by_count(const uint32_t* __restrict__ drive, uint32_t ncol, uint32_t inCols,
                                              uint32_t target, uint32_t dv) {
    // lane j < WG: wave j's column
    const uint32_t dl = drive[min(emit_col<WG>(blockIdx.x, lane_id() & (WG - 1u)), ncol - 1u)];
    const uint32_t above = drives_above<WG, 64u * WG>(drive, ncol, inCols, dl, threadIdx.x >> 6);  // a barrier inside
    return dv >= 1u && above < target;
}

// Column gate and emit in one kernel for a single network, one wave per
// column, 16 columns per workgroup.  In the usual case (gate_by_count) each
// workgroup counts drives above its columns' instead of selecting theta;
// theta itself is then computed only where it is an output (the side
// statistics of a synchronous presentation: workgroup 0 runs the radix
// select and stores (theta, need)).  Otherwise every workgroup selects theta
// over all column drives itself (k_topk_par / k_topk_serial, mm:259-352) --
// LDS work on 16 KiB the drive kernel left in cache -- instead of waiting for
// a separate one-workgroup launch.
// Every wave of the workgroup calls it: the gate ends with a workgroup
// barrier.  Returns the column's spike word and margin flag.
template <bool LEARN, bool STATS, uint32_t WG = 16>
__device__ __forceinline__ uint2 emit_step(const NetDev& d, const Params& P, const uint32_t* __restrict__ spkIn,
                                           uint32_t* __restrict__ spkOut, uint32_t col) {
    EmitIn in = emit_inputs<LEARN, STATS>(d, P, min(col, d.ncol - 1u));  // in flight during the gate
    const uint32_t nrec = d.ncol - P.inCols;
    bool gate;
    if (P.targetCols >= 1u && P.targetCols <= nrec && nrec <= 4096u) {  // uniform over the grid
        if (STATS && blockIdx.x == 0u)
            (void)topk_block<64u * WG>(d.drive, d.theta, d.ncol, P.inCols, P.targetCols, d.topkpar);
        gate = gate_by_count<WG>(d.drive, d.ncol, P.inCols, P.targetCols, in.dv);
    } else {
        const uint32_t th = topk_block<64u * WG>(d.drive, blockIdx.x == 0u ? d.theta : nullptr, d.ncol, P.inCols,
                                                 P.targetCols, d.topkpar);
        gate = in.dv >= th;
    }
    if (col < d.ncol && (gate || (P.clamp != 0u && col < P.inCols)))  // wave-uniform: this column can spike
        emit_inputs_spiking<LEARN, STATS>(d, P, col, in);
    return emit_cols<LEARN, STATS>(d, P, spkIn, spkOut, col, in, gate);
}
// WG columns (waves) per workgroup: 16, or 8 for the training emit
// (EMIT_TRAIN_WAVES).
template <bool LEARN, bool STATS, uint32_t WG = 16>
__global__ __launch_bounds__(64 * WG) void k_emit(NetDev d, Params P, const uint32_t* __restrict__ spkIn,
                                                  uint32_t* __restrict__ spkOut) {
    (void)emit_step<LEARN, STATS, WG>(d, P, spkIn, spkOut, emit_col<WG>(blockIdx.x, threadIdx.x >> 6));
}
// Measured on one 131072-neuron network with 16-column propagation: the
// training emit in 4-column workgroups 15.65k GSyn/s, 8 columns 16.40k; the
// last step's emit with plasticity in 8-column workgroups 16.42k, 16 columns
// 16.40k (the same within run-to-run spread).  With 4-column propagation the
// training emit went from 15.75k (16 columns) to 16.05k (8 columns).
constexpr uint32_t EMIT_TRAIN_WAVES = 8;
constexpr uint32_t EMIT_LEARN_WAVES = 16;

// Columns from an emit workgroup's first to its last (emit_col: WG /
// DRIVE_COLS runs of DRIVE_COLS columns, 8 * DRIVE_COLS apart, or WG
// consecutive columns when the emit workgroup takes part of a propagation
// workgroup).
template <uint32_t WG>
constexpr uint32_t emit_span() { return WG >= DRIVE_COLS ? DRIVE_COLS * 8u * (WG / DRIVE_COLS - 1u) + DRIVE_COLS : WG; }

// The last timestep of a presentation without side statistics, when it
// learns: the emit of k_emit<true, false>, then that step's plasticity for
// the workgroup's own columns in the same workgroup.  A column's plasticity
// needs its own spike word and margin (just emitted here), its own NMDA bits,
// fired masks and band flags (read with the emit's inputs) and the step's
// presynaptic spikes, so no launch of its own ends the presentation.  Its
// work is counted in cntLast.
template <uint32_t WG>
__global__ __launch_bounds__(64 * WG) void k_emit_learn(NetDev d, Params P, const uint32_t* __restrict__ spkIn,
                                                        uint32_t* __restrict__ spkOut) {
    __shared__ VoteList<WG> vl;
    const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
    const uint32_t col = emit_col<WG>(blockIdx.x, wave), ccol = min(col, d.ncol - 1u);
    const uint32_t gl = d.conn[ccol * NSLOT + (lane & 31u)];
    const VoteIn vi = vote_inputs(d, ccol, reinterpret_cast<const uint2*>(d.cntLast));
    if (threadIdx.x == 0u) { vl.nt = 0u; vl.nl = 0u; }
    __syncthreads();
    const uint2 bm = emit_step<true, false, WG>(d, P, spkIn, spkOut, col);
    if (col < d.ncol && lane < 32u && learns(P, bm.x, bm.y))
        vote_column<WG>(d, P, col, bm.x, gl, vi, 0u, reinterpret_cast<uint2*>(d.cntLast), vl);
    __syncthreads();
    if (vl.nl != 0u)  // workgroup-uniform
        rewrite_voted<WG, 4, emit_span<WG>(), ROUNDS>(d, P, spkIn, vl, wave, WG, emit_col<WG>(blockIdx.x, 0u) * NSLOT);
}

// The first timestep of a presentation without side statistics.  Its
// propagation would read the cleared spikes: every accumulator, NMDA bit
// and drive is 0, so no column passes the gate (drive >= 1) and only clamped
// input columns spike, with their stimulus word and no voted tile (no NMDA
// bits).  This kernel emits that step directly -- spikes, history row,
// margin and spiking-column flags, the bias update -- in place of the clear,
// the propagation and the (empty) plasticity launches.  The previous step's
// spikes are not read.
template <bool LEARN>
__global__ __launch_bounds__(1024) void k_emit_first(NetDev d, Params P, uint32_t* __restrict__ spkOut) {
    const uint32_t col = emit_col(blockIdx.x, threadIdx.x >> 6);
    if (col >= d.ncol) return;
    const uint32_t lane = lane_id(), n = lane & 31u;
    const uint32_t gid = col * NPC + n;
    const bool rec = col >= P.inCols;
    const uint32_t b = (P.clamp != 0u && col < P.inCols) ? d.stim[col] : 0u;
    if (lane == 0u) {
        spkOut[col] = b;
        d.marg[col] = 0u;  // all 32 accumulators are 0
        d.colNZ[col] = b != 0u ? 1u : 0u;
        if (P.histOn != 0u) d.hist[(size_t)P.step * d.ncol + col] = b;
    }
    if (LEARN && lane < 32u && rec) {  // k_bias (recurrent columns never spike here)
        const int dec = ((P.step % max(P.biasDownEvery, 1u)) == 0u) ? (int)P.biasDown : 0;
        d.bias[gid] = (uint32_t)max(0, (int)d.bias[gid] - dec);
    }
}

template <bool STATS>
__global__ __launch_bounds__(256) void k_learn(NetDev d, Params P, const uint32_t* __restrict__ spkIn,
                                               const uint32_t* __restrict__ spkOut) {
    learn_cols<STATS>(d, P, spkIn, spkOut);
}

// ============================================================== readouts
// k_wstat, mm:641-646: functional synapse count (popcount of W3)
__global__ __launch_bounds__(256) void k_wstat(const uint4* __restrict__ W3, size_t n, uint32_t* out) {  // any layout
    uint32_t s = 0;
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256ull) {
        const uint4 w = W3[i];
        s += popc(w.x) + popc(w.y) + popc(w.z) + popc(w.w);
    }
    s = wave_sum(s);
    if (lane_id() == 0u) atomicAdd(out, s);
}

// ---- readout reductions over the word-major planes (one wave per column) ----
// A (column, branch) tile of a plane is 256 words = 1 KiB, and lane L reads
// words 4L..4L+3 of it with one dwordx4: presynaptic word w = L / 8 (column
// 8g + w of the branch's group g) of neurons 4 (L % 8) .. 4 (L % 8) + 3.  A
// column's tiles are contiguous, so a wave streams them as 1 KiB coalesced
// requests.  Column totals are summed over the wave, per-neuron totals over
// the eight lanes that hold the same neurons; every sum is u32 wrap-around,
// so the order of additions does not change a result.
__device__ __forceinline__ uint4 tile_quad(const uint32_t* __restrict__ plane, uint32_t tile) {
    return reinterpret_cast<const uint4*>(plane + (size_t)tile * TILE_WORDS)[lane_id()];
}
// Sum over the eight lanes with the same lane % 8 (lane bits 3-5).
__device__ __forceinline__ uint32_t octet_sum(uint32_t x) {
    x += __shfl_xor(x, 8);
    x += __shfl_xor(x, 16);
    x += __shfl_xor(x, 32);
    return x;
}
// Lane i (< NV) adds counter i of the wave's totals to out[i]: one atomic
// instruction per wave.
template <int NV>
__device__ __forceinline__ void wave_add_counters(const uint32_t (&v)[NV], uint32_t* __restrict__ out) {
    const uint32_t lane = lane_id();
    uint32_t mine = 0u;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const uint32_t t = wave_sum(v[i]);
        if (lane == (uint32_t)i) mine = t;
    }
    if (lane < (uint32_t)NV && mine != 0u) atomicAdd(&out[lane], mine);
}

// Per-neuron counts over the recurrent branches (s >= inbr), for every
// neuron: KIND 0 functional synapses, the popcount of W3 (k_indeg,
// mm:566-581); KIND 1 synapses still at an initial counter value 5 or 6,
// i.e. W3 = 0, W2 = 1 and W1 != W0 (k_untouched, mm:501-519).  Four tiles'
// loads are in flight per wave.
template <int KIND>
__global__ __launch_bounds__(256) void k_neuron_counts(NetDev d, uint32_t* __restrict__ out) {
    const uint32_t col = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (col >= d.ncol) return;  // wave-uniform
    constexpr uint32_t U = 4;
    constexpr int NP = KIND == 0 ? 1 : 4;
    uint32_t c[4] = {0u, 0u, 0u, 0u};
    for (uint32_t s0 = d.inbr; s0 < NSLOT; s0 += U) {
        uint4 a[U][NP];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const uint32_t tile = col * NSLOT + min(s0 + u, NSLOT - 1u);  // past the end: re-read, not counted
            a[u][0] = tile_quad(d.W3, tile);
            if (KIND == 1) {
                a[u][1] = tile_quad(d.W2, tile);
                a[u][2] = tile_quad(d.W1, tile);
                a[u][3] = tile_quad(d.W0, tile);
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            if (s0 + u >= NSLOT) break;
            const uint32_t x3[4] = {a[u][0].x, a[u][0].y, a[u][0].z, a[u][0].w};
            if (KIND == 0) {
#pragma unroll
                for (int j = 0; j < 4; ++j) c[j] += popc(x3[j]);
            } else {
                const uint32_t x2[4] = {a[u][1].x, a[u][1].y, a[u][1].z, a[u][1].w};
                const uint32_t x1[4] = {a[u][2].x, a[u][2].y, a[u][2].z, a[u][2].w};
                const uint32_t x0[4] = {a[u][3].x, a[u][3].y, a[u][3].z, a[u][3].w};
#pragma unroll
                for (int j = 0; j < 4; ++j) c[j] += popc(~x3[j] & x2[j] & (x1[j] ^ x0[j]));
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) c[j] = octet_sum(c[j]);
    const uint32_t lane = lane_id();
    if (lane < 8u) reinterpret_cast<uint4*>(out + (size_t)col * NPC)[lane] = make_uint4(c[0], c[1], c[2], c[3]);
}

// Functional recurrent density from one assembly onto selected postsynaptic
// neurons (k_within, mm:524-562).  Selected: neurons of recurrent columns
// whose postMask word is nonzero (useMask) or whose bit is set in `ball`.
// For each selected neuron and recurrent branch, the branch's 256 presynaptic
// bits are split by the ballot of their column in `ball`: in (set) and out.
// out[0..3] = functional in, all in, functional out, all out; out[4], out[5] =
// functional / all in for selected neurons shared with another assembly
// (mult > 1), out[6], out[7] the same for the others; out[8], out[9] = the
// numbers of shared / other selected neurons.  A column with no selected
// neuron costs one ballot.
__global__ __launch_bounds__(256) void k_within(NetDev d, const uint32_t* __restrict__ ball,
                                                const uint32_t* __restrict__ postMask, uint32_t useMask,
                                                uint32_t* __restrict__ out) {
    const uint32_t col = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (col >= d.ncol || col < d.incols) return;  // wave-uniform
    const uint32_t lane = lane_id(), n = lane & 31u;
    const uint32_t gid = col * NPC + n;
    const uint32_t sel = useMask ? ballot32(postMask[gid] != 0u) : ball[col];
    if (sel == 0u) return;
    const uint32_t sh = ballot32(d.mult[gid] > 1u);
    const uint32_t cn = d.conn[col * NSLOT + n];
    const uint32_t q0 = (lane & 7u) * 4u, w = lane >> 3;
    const uint32_t selq = (sel >> q0) & 0xFu, shq = (sh >> q0) & 0xFu;
    uint32_t fS = 0, pS = 0, fE = 0, pE = 0, fO = 0, pO = 0;
    constexpr uint32_t U = 4;  // tiles in flight
    for (uint32_t s0 = d.inbr; s0 < NSLOT; s0 += U) {
        uint32_t ap[U];
        uint4 x[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const uint32_t s = min(s0 + u, NSLOT - 1u);  // past the end: re-read, not counted
            ap[u] = ball[8u * __builtin_amdgcn_readlane(cn, s) + w];
            x[u] = tile_quad(d.W3, col * NSLOT + s);
        }
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            if (s0 + u >= NSLOT) break;
            const uint32_t xv[4] = {x[u].x, x[u].y, x[u].z, x[u].w};
            const uint32_t pa = popc(ap[u]);
#pragma unroll
            for (uint32_t j = 0; j < 4u; ++j) {
                if (((selq >> j) & 1u) == 0u) continue;
                const uint32_t fi = popc(xv[j] & ap[u]);
                if ((shq >> j) & 1u) { fS += fi; pS += pa; } else { fE += fi; pE += pa; }
                fO += popc(xv[j] & ~ap[u]);
                pO += 32u - pa;
            }
        }
    }
    // the ballots are wave-uniform: their counts enter the wave sum from lane 0 only
    const uint32_t nS = lane == 0u ? popc(sel & sh) : 0u, nE = lane == 0u ? popc(sel & ~sh) : 0u;
    const uint32_t v[10] = {fS + fE, pS + pE, fO, pO, fS, pS, fE, pE, nS, nE};
    wave_add_counters<10>(v, out);
}

// OR of the assemblies' ballots of each column (the presynaptic columns that
// are members of some assembly), for k_wdiag.
__global__ __launch_bounds__(256) void k_asm_union(const uint32_t* __restrict__ asmSpk, uint32_t nstim, uint32_t ncol,
                                                   uint32_t* __restrict__ any) {
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= ncol) return;
    uint32_t o = 0u;
    for (uint32_t q = 0; q < nstim; ++q) o |= asmSpk[(size_t)q * ncol + c];
    any[c] = o;
}

// Counter-value histogram split by the assembly relation of the synapse's two
// neurons (k_wdiag, mm:591-638), over postsynaptic neurons that belong to an
// assembly (memb bit q: member of assembly q < nstim) and over the branches
// `sel` names: 0 recurrent, 1 afferent, otherwise all.  A synapse from
// presynaptic neuron x is "within" when x is in the ballot of an assembly the
// neuron belongs to, "between" when x is in some assembly's ballot but none
// of those.  hist[v], hist[16 + v], hist[32 + v]: synapses with counter value
// v that are within, between, and all.  The sixteen value masks of a word are
// built as a decision tree over the four planes (W3 first); a neuron's
// "within" word ORs only the ballots of its own assemblies (few per neuron).
// `any` is k_asm_union's output.
__global__ __launch_bounds__(256) void k_wdiag(NetDev d, const uint32_t* __restrict__ memb,
                                               const uint32_t* __restrict__ asmSpk,
                                               const uint32_t* __restrict__ any, uint32_t nstim, uint32_t sel,
                                               uint32_t* __restrict__ hist) {
    const uint32_t col = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (col >= d.ncol) return;  // wave-uniform
    const uint32_t lane = lane_id(), n = lane & 31u;
    // a neuron takes part when its membership word is nonzero; its "within"
    // word ORs the ballots of its assemblies below nstim
    const uint32_t qmask = nstim >= 32u ? 0xFFFFFFFFu : ((1u << nstim) - 1u);
    if (ballot32(memb[col * NPC + n] != 0u) == 0u) return;
    const uint4 mq4 = reinterpret_cast<const uint4*>(memb + (size_t)col * NPC)[lane & 7u];
    const uint32_t mq[4] = {mq4.x, mq4.y, mq4.z, mq4.w};
    const uint32_t cn = d.conn[col * NSLOT + n];
    const uint32_t w = lane >> 3;
    const uint32_t sBeg = sel == 0u ? d.inbr : 0u, sEnd = sel == 1u ? d.inbr : NSLOT;
    uint32_t acc[48];
#pragma unroll
    for (int i = 0; i < 48; ++i) acc[i] = 0u;
    for (uint32_t s = sBeg; s < sEnd; ++s) {
        const uint32_t c = 8u * __builtin_amdgcn_readlane(cn, s) + w;
        const uint32_t tile = col * NSLOT + s;
        const uint4 A3 = tile_quad(d.W3, tile), A2 = tile_quad(d.W2, tile), A1 = tile_quad(d.W1, tile),
                    A0 = tile_quad(d.W0, tile);
        const uint32_t anyp = any[c];
        const uint32_t a3[4] = {A3.x, A3.y, A3.z, A3.w}, a2[4] = {A2.x, A2.y, A2.z, A2.w},
                       a1[4] = {A1.x, A1.y, A1.z, A1.w}, a0[4] = {A0.x, A0.y, A0.z, A0.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (mq[j] == 0u) continue;
            uint32_t within = 0u;
            for (uint32_t m = mq[j] & qmask; m; m &= m - 1u)
                within |= asmSpk[(size_t)__builtin_ctz(m) * d.ncol + c];
            const uint32_t between = anyp & ~within;
            uint32_t e[16];  // e[v]: the synapses whose counter is v
            e[0] = ~a3[j];
            e[8] = a3[j];
#pragma unroll
            for (int b = 2; b >= 0; --b) {  // split every mask on plane b
                const uint32_t pb = b == 2 ? a2[j] : b == 1 ? a1[j] : a0[j];
                const int step = 1 << (b + 1);
#pragma unroll
                for (int v = 0; v < 16; v += step) {
                    e[v + (step >> 1)] = e[v] & pb;
                    e[v] &= ~pb;
                }
            }
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                acc[v] += popc(e[v] & within);
                acc[16 + v] += popc(e[v] & between);
                acc[32 + v] += popc(e[v]);
            }
        }
    }
    wave_add_counters<48>(acc, hist);
}

// k_bstest, mm:424-434
__global__ __launch_bounds__(256) void k_bstest(uint32_t op, uint32_t lo, uint32_t hi, uint32_t* a0, uint32_t* a1,
                                                uint32_t* a2, uint32_t* a3, const uint32_t* __restrict__ m,
                                                size_t n) {
    const size_t i = blockIdx.x * 256ull + threadIdx.x;
    if (i >= n) return;
    uint32_t x0 = a0[i], x1 = a1[i], x2 = a2[i], x3 = a3[i];
    if (op == 0u) bs_inc(x0, x1, x2, x3, m[i]);
    else if (op == 1u) bs_dec(x0, x1, x2, x3, m[i]);
    else bs_clip(x0, x1, x2, x3, lo, hi);
    a0[i] = x0; a1[i] = x1; a2[i] = x2; a3[i] = x3;
}

__global__ __launch_bounds__(256) void k_kwta_test(const uint32_t* __restrict__ v, uint32_t ncols, uint32_t k,
                                                   uint32_t* ball, uint32_t* vk, uint32_t* vk1) {
    const uint32_t col = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (col >= ncols) return;
    const uint32_t n = lane_id() & 31u;
    const Kwta r = kwta32(v[(size_t)col * 32u + n], k, n);
    if (lane_id() == 0u) { ball[col] = r.ball; vk[col] = r.vk; vk1[col] = r.vk1; }
}

// ============================================================== probe batches
// Many recall probes against one resident network (readout presentations:
// only W3, conn and optionally bias are read, nothing persistent is written).
// Each probe slot keeps its own spikes, drives, accumulators and history.
//
// k_probe_drive<PT>: one wave per (column, probe tile of PT slots).  For each
// branch, the W3 tile is loaded once into registers and reused by every probe
// of the tile whose presynaptic group fired; per-probe accumulators live in
// registers (PT of them), and the probe loop is unrolled so a probe whose
// group is silent costs one scalar branch.  Group activity comes from a bit
// per presynaptic group (gact) that k_probe_select writes with the spikes.
// DB branch tiles (and all their probes' spike words) are in flight together;
// the per-slot state is loaded in one batch before the branch loop and after
// it, so a wave waits for memory a few times, not once per slot.
//
// Blocks are numbered so that the blocks one XCD receives (linear id L,
// L+8, L+16, ...) walk all probe tiles of the same four columns back to back,
// so the tile's W3 bytes are reused from that XCD's L2.
// Probe propagation for tiles of up to 8 probes, as drive_col does it for one
// network: work items (branch, spiking word, probe) listed in LDS per
// half-wave (all of a branch's items, for every probe, go to one half;
// branches with work are dealt alternately to the halves), so one dword load
// instruction serves one item of each half and 8 items per half are in
// flight together.  An item reads the W3 word of (branch, word) for neuron n
// and the probe's spike word of presynaptic column 8g + w.  Per wave this
// takes a few round trips instead of one per branch of the tile's union
// (the probes of a tile share few W3 lines, so that union saves little).
//
// The kernel is bound by instruction issue, so an item carries its load
// offsets ready-made:
//   bits 0-7   branch * 8 + word     (W3 word index within the column tile)
//   bits 8-10  probe within the tile
//   bits 11-30 probe * ncol + 8g + word  (spike word index; needs PT*ncol <= 2^20)
//   bit  31    last item of its (probe, branch) pair (a sign test)
// Each half's list is padded to a whole round with PAD items (spike index out
// of range: a zero word; never last), so a round has no per-item bounds test.
// A (probe, branch) count is thresholded when its last item is consumed and
// added to the neuron's accumulator of that probe in LDS (one per probe and
// lane, shared by both halves).
//
// A wave walks `tpw` consecutive tiles of its column: conn and bias are read
// once, the next tile's group-activity words are requested before this
// tile's items, and each tile's slot state is requested with its activity
// words.
constexpr uint32_t PROBE_PAD_ITEM = 0xFFFFFu << 11;
// K items of a half's list (K = 8 or 4): their W3 and spike words are
// requested together, then each count is added up and, on a pair's last
// item, thresholded into the neuron's accumulator of that probe (An = A + n).
template <uint32_t K, int MODE>
__device__ __forceinline__ void probe_round(const uint32_t* Li, __amdgpu_buffer_rsrc_t w3, __amdgpu_buffer_rsrc_t sp,
                                            uint32_t n4, uint32_t* An, uint32_t& c, uint32_t dthr, uint32_t boost) {
    uint32_t it[K];
    const uint4 i0 = *reinterpret_cast<const uint4*>(Li);
    it[0] = i0.x; it[1] = i0.y; it[2] = i0.z; it[3] = i0.w;
    if constexpr (K == 8) {
        const uint4 i1 = *reinterpret_cast<const uint4*>(Li + 4u);
        it[4] = i1.x; it[5] = i1.y; it[6] = i1.z; it[7] = i1.w;
    }
    uint32_t wv[K], sv[K];
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
        wv[k] = buf_ld(w3, ((it[k] & 0xFFu) << 7) + n4);
        sv[k] = buf_ld(sp, ((it[k] >> 11) & 0xFFFFFu) << 2);
    }
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
        const uint32_t e = it[k];
        c += popc(wv[k] & sv[k]);
        if ((int32_t)e < 0) {  // last item of this (probe, branch): its count is complete
            // mode 0: the count; mode 1: + boost at the threshold; otherwise 1 at the threshold
            const uint32_t x = MODE == 0 ? c : MODE == 1 ? (c >= dthr ? c + boost : c) : (c >= dthr ? 1u : 0u);
            atomicAdd(&An[((e >> 8) & 7u) * 32u], x);
            c = 0u;
        }
    }
}
template <int PT, int MODE>
__global__ __launch_bounds__(256, PT > 4 ? 4 : 8) void k_probe_items(ProbeDev d, uint32_t ncolgroups, uint32_t ntiles, uint32_t tpw) {
    static_assert(PT == 1 || PT == 2 || PT == 4 || PT == 8, "three probe bits in an item");
    using Mask = typename std::conditional<(PT > 4), uint64_t, uint32_t>::type;
    // a half's list: at most (total + one branch's items) / 2 of the 32 x 8 x PT items, padded to 4
    constexpr uint32_t PROBE_ITEMS_HALF = 132u * PT;
    constexpr uint32_t DB = 8;  // items per half in flight
    __shared__ __align__(16) uint32_t items[4][2 * PROBE_ITEMS_HALF];
    __shared__ uint32_t accs[4][PT * 32];
    const uint32_t nchunks = (ntiles + tpw - 1u) / tpw;
    const uint32_t L = blockIdx.x;
    const uint32_t q = L >> 3;
    const uint32_t chunk = q % nchunks;
    const uint32_t cg = (q / nchunks) * 8u + (L & 7u);
    if (cg >= ncolgroups) return;
    const uint32_t lane = lane_id(), n = lane & 31u;
    const bool hi = lane >= 32u;
    const uint32_t col = cg * 4u + (threadIdx.x >> 6);
    if (col >= d.ncol) return;
    const uint32_t base = col * NSLOT;
    const uint32_t gid = base + n;
    const uint32_t gl = d.conn[gid];
    const uint32_t bi = d.bias[gid] >> d.biasShift;
    const uint32_t t0 = chunk * tpw, t1 = min(t0 + tpw, ntiles);
    const uint32_t ubase = __builtin_amdgcn_readfirstlane(base);
    const __amdgpu_buffer_rsrc_t w3 = buf_rsrc(d.W3 + (size_t)ubase * TILE_WORDS, NSLOT * TILE_WORDS * 4u);
    uint32_t* W = items[threadIdx.x >> 6];
    uint32_t* A = accs[threadIdx.x >> 6];
    uint32_t* Li = W + (hi ? PROBE_ITEMS_HALF : 0u);
    // per-slot state and activity words through range-checked resources (32-bit offsets)
    const __amdgpu_buffer_rsrc_t ga = buf_rsrc(d.gact, d.nslots * d.gw * 4u);
    const __amdgpu_buffer_rsrc_t rt = buf_rsrc(d.st_t, d.nslots * 4u), rT = buf_rsrc(d.st_T, d.nslots * 4u);
    const __amdgpu_buffer_rsrc_t rb = buf_rsrc(d.st_bias, d.nslots * 4u);
    const uint32_t gaOff = (gl >> 2) * 4u;
    uint32_t gaw[PT];
#pragma unroll
    for (int p = 0; p < PT; ++p)  // slots past the end read the last slot's words and are masked off
        gaw[p] = buf_ld(ga, min(t0 * PT + (uint32_t)p, d.nslots - 1u) * d.gw * 4u + gaOff);
    for (uint32_t tile = t0; tile < t1; ++tile) {
        const uint32_t slot0 = tile * PT;
        // slot state of the tile's probes, one probe per lane, as bit masks
        const uint32_t slp = slot0 + n;
        const bool slv = n < (uint32_t)PT && slp < d.nslots;
        const uint32_t slc = min(slp, d.nslots - 1u);
        const uint32_t st_t = buf_ld(rt, slc * 4u), st_T = buf_ld(rT, slc * 4u), st_b = buf_ld(rb, slc * 4u);
        uint32_t fs[PT];
#pragma unroll
        for (int p = 0; p < PT; ++p)
            fs[p] = d.facil != 0u ? d.spk[(size_t)min(slot0 + (uint32_t)p, d.nslots - 1u) * d.ncol + col] : 0u;
        const uint32_t live = ballot32(slv && st_t < st_T);
        const uint32_t ubm = ballot32(slv && st_b != 0u);
        Mask fm = 0;  // bit 8p + w: presynaptic column 8g + w of this lane's branch spiked in probe p
#pragma unroll
        for (int p = 0; p < PT; ++p)
            if ((live >> p) & 1u) fm |= (Mask)((gaw[p] >> (8u * (gl & 3u))) & 0xFFu) << (8 * p);
        if (tile + 1u < t1) {
#pragma unroll
            for (int p = 0; p < PT; ++p)
                gaw[p] = buf_ld(ga, min(slot0 + PT + (uint32_t)p, d.nslots - 1u) * d.gw * 4u + gaOff);
        }
        if (live == 0u) continue;
#pragma unroll
        for (uint32_t i = lane; i < PT * 32u; i += 64u) A[i] = 0u;
        const uint32_t ni = sizeof(Mask) == 8 ? (uint32_t)__popcll((unsigned long long)fm) : popc((uint32_t)fm);
        // Branches go to the halves in lane order, split where the tile's items halve, so the
        // halves' lists differ by less than one branch's items (both halves compute the same scan).
        const uint32_t sc = half_scan(ni);
        const uint32_t total = __builtin_amdgcn_readlane(sc, 31);
        const uint32_t ex = sc - ni;
        const uint32_t hh = (2u * ex + ni > total) ? 1u : 0u;  // non-decreasing in lane order
        const uint32_t b1 = ballot32(hh != 0u);
        const uint32_t nA = b1 != 0u ? __builtin_amdgcn_readlane(ex, __builtin_ctz(b1)) : total;
        const uint32_t nB = total - nA;
        {
            // lane n lists branch n's items of the first PH probes, lane n + 32 those of the others
            constexpr int PH = PT >= 2 ? PT / 2 : PT;
            const Mask lowm = PT >= 2 ? (Mask)(((Mask)1 << (8 * PH)) - 1u) : (Mask)~(Mask)0;
            const Mask mine = hi ? (PT >= 2 ? (Mask)(fm & ~lowm) : (Mask)0) : (Mask)(fm & lowm);
            uint32_t* Lh = W + hh * PROBE_ITEMS_HALF;
            const uint32_t nlow = sizeof(Mask) == 8 ? (uint32_t)__popcll((unsigned long long)(fm & lowm))
                                                    : popc((uint32_t)(fm & lowm));
            uint32_t pos = (hh ? ex - nA : ex) + (hi ? nlow : 0u);
            const uint32_t e0 = (n << 3) | ((8u * gl) << 11);  // branch and group parts
            Mask m = mine;
            while (m) {
                const uint32_t bit = sizeof(Mask) == 8 ? (uint32_t)__builtin_ctzll((unsigned long long)m)
                                                       : (uint32_t)__builtin_ctz((uint32_t)m);
                m &= m - 1u;
                const uint32_t pp = bit >> 3, w = bit & 7u;
                const bool last = ((m >> (bit & ~7u)) & (Mask)0xFFu) == 0u;
                Lh[pos++] = e0 + w + (pp << 8) + (last ? 0x80000000u : 0u) + ((pp * d.ncol + w) << 11);
            }
        }
        const uint32_t myN = hi ? nB : nA;
        const uint32_t R = (max(nA, nB) + 3u) & ~3u;  // rounds of DB items, then one of 4 if needed
        for (uint32_t i = myN + n; i < R; i += 32u) Li[i] = PROBE_PAD_ITEM;
        const uint32_t nsl = min((uint32_t)PT, d.nslots - slot0);  // slots of this tile that exist
        const __amdgpu_buffer_rsrc_t sp = buf_rsrc(d.spk + (size_t)slot0 * d.ncol, nsl * d.ncol * 4u);
        const uint32_t n4 = n * 4u;
        uint32_t c = 0u;
        uint32_t r = 0;
        for (; r + DB <= R; r += DB) probe_round<DB, MODE>(Li + r, w3, sp, n4, A + n, c, d.dthr, d.boost);
        if (r < R) probe_round<4, MODE>(Li + r, w3, sp, n4, A + n, c, d.dthr, d.boost);
        // Epilogue: each half-wave finishes its own probes (the low half the
        // first PE, the high half the rest), so one scan serves two probes.
        constexpr int PE = PT >= 2 ? PT / 2 : 1;  // probes per half
        const uint32_t pbase = (PT >= 2 && hi) ? (uint32_t)PE : 0u;
        const bool halfOn = PT >= 2 || !hi;
        uint32_t av[PE];
#pragma unroll
        for (int q = 0; q < PE; ++q) av[q] = A[(pbase + (uint32_t)q) * 32u + n];
        const uint32_t aw = d.acc16 ? 2u : 4u;  // bytes per accumulator
        const __amdgpu_buffer_rsrc_t ar = buf_rsrc(reinterpret_cast<const uint8_t*>(d.acc) +
                                                       (size_t)slot0 * d.ncol * NPC * aw,
                                                   nsl * d.ncol * NPC * aw);
        const __amdgpu_buffer_rsrc_t dr = buf_rsrc(d.drive + (size_t)slot0 * d.ncol, nsl * d.ncol * 4u);
#pragma unroll
        for (int q = 0; q < PE; ++q) {
            const uint32_t p = pbase + (uint32_t)q;
            const bool lv = halfOn && ((live >> p) & 1u) != 0u;
            if (__ballot(lv) == 0ull) continue;  // neither half's probe is live
            uint32_t fsp;
            if constexpr (PT >= 2) fsp = hi ? fs[q + PE] : fs[q];
            else fsp = fs[0];
            uint32_t a = av[q];
            if (((fsp >> n) & 1u) != 0u) a += d.facil;
            if (((ubm >> p) & 1u) != 0u) a = (a > bi) ? (a - bi) : 0u;
            const uint32_t ao = lv ? (p * d.ncol * NPC + gid) * aw : SKIP;
            if (d.acc16) buf_st16(ar, ao, a);
            else buf_st(ar, ao, a);
            const uint32_t drv = half_scan(a);
            buf_st(dr, (lv && n == 31u) ? (p * d.ncol + col) * 4u : SKIP, drv);
        }
    }
}

template <int PT>
__global__ __launch_bounds__(256) void k_probe_drive(ProbeDev d, uint32_t ncolgroups, uint32_t ntiles) {
    constexpr int DB = 1;                // branch tiles in flight together
    constexpr int PC = PT < 8 ? PT : 8;  // probes whose spike words are in flight together
    const uint32_t L = blockIdx.x;
    const uint32_t q = L >> 3;
    const uint32_t tile = q % ntiles;
    const uint32_t cg = (q / ntiles) * 8u + (L & 7u);
    if (cg >= ncolgroups) return;
    const uint32_t lane = lane_id(), n = lane & 31u;
    const bool hi = lane >= 32u;
    const uint32_t col = cg * 4u + (threadIdx.x >> 6);
    if (col >= d.ncol) return;
    const uint32_t base = col * NSLOT;
    const uint32_t slot0 = tile * PT;
    const uint32_t gl = d.conn[base + n];

    // slots past the end read the last slot's state and are masked off
    // gact holds one bit per column: the byte of group gl is bits 8gl..8gl+7
    uint32_t stt[PT], stT[PT], gaw[PT];
#pragma unroll
    for (int p = 0; p < PT; ++p) {
        const uint32_t sl = min(slot0 + (uint32_t)p, d.nslots - 1u);
        stt[p] = d.st_t[sl];
        stT[p] = d.st_T[sl];
        gaw[p] = d.gact[(size_t)sl * d.gw + (gl >> 2)];
    }
    uint32_t bm[PT];
    uint32_t anyb = 0, fwu = 0;  // fwu (lane s < 32): columns of slot s's group that spiked in any live probe
    uint64_t live = 0;
#pragma unroll
    for (int p = 0; p < PT; ++p) {
        const bool lv = slot0 + (uint32_t)p < d.nslots && stt[p] < stT[p];
        live |= (uint64_t)(lv ? 1u : 0u) << p;
        const uint32_t fb = lv ? (gaw[p] >> (8u * (gl & 3u))) & 0xFFu : 0u;
        fwu |= fb;
        bm[p] = ballot32(fb != 0u);
        anyb |= bm[p];
    }
    if (live == 0u) return;

    uint32_t acc[PT];
#pragma unroll
    for (int p = 0; p < PT; ++p) acc[p] = 0u;
    const uint32_t ubase = __builtin_amdgcn_readfirstlane(base);
    const __amdgpu_buffer_rsrc_t w3 = buf_rsrc(d.W3 + (size_t)ubase * TILE_WORDS, NSLOT * TILE_WORDS * 4u);
    uint32_t mask = anyb;
    while (mask) {
        uint32_t sidx[DB], g[DB], fw[DB];
#pragma unroll
        for (int k = 0; k < DB; ++k) {
            sidx[k] = 32u; g[k] = 0u; fw[k] = 0u;
            if (mask) {
                const uint32_t sk = __builtin_ctz(mask);
                mask &= mask - 1u;
                sidx[k] = sk;
                fw[k] = __builtin_amdgcn_readlane(fwu, sk);
                g[k] = __builtin_amdgcn_readlane(gl, sk);
            }
        }
        uint4 w[DB];
#pragma unroll
        for (int k = 0; k < DB; ++k) {  // the W3 lines of columns that spiked in some probe of the tile
            const uint32_t my = (fw[k] >> (hi ? 4u : 0u)) & 0xFu;
            const uint32_t o = ((sidx[k] * 8u + (hi ? 4u : 0u)) * 32u + n) * 4u;
            w[k].x = buf_ld(w3, (my & 1u) ? o : SKIP);
            w[k].y = buf_ld(w3, (my & 2u) ? o + 128u : SKIP);
            w[k].z = buf_ld(w3, (my & 4u) ? o + 256u : SKIP);
            w[k].w = buf_ld(w3, (my & 8u) ? o + 384u : SKIP);
        }
#pragma unroll
        for (int p0 = 0; p0 < PT; p0 += PC) {
            uint4 sv[DB][PC];
#pragma unroll
            for (int k = 0; k < DB; ++k) {
                if (sidx[k] >= 32u) continue;
#pragma unroll
                for (int pp = 0; pp < PC; ++pp)
                    if ((bm[p0 + pp] >> sidx[k]) & 1u)
                        sv[k][pp] = group_words(d.spk + (size_t)(slot0 + p0 + pp) * d.ncol, g[k], hi);
            }
#pragma unroll
            for (int k = 0; k < DB; ++k) {
                if (sidx[k] >= 32u) continue;
#pragma unroll
                for (int pp = 0; pp < PC; ++pp) {
                    const int p = p0 + pp;
                    if ((bm[p] >> sidx[k]) & 1u) {
                        const uint4 v = sv[k][pp];
                        uint32_t c = popc(w[k].x & v.x) + popc(w[k].y & v.y) + popc(w[k].z & v.z) +
                                     popc(w[k].w & v.w);
                        c += __shfl_xor(c, 32);
                        const bool nmda = c >= d.dthr;
                        if (d.mode == 0u) acc[p] += c;
                        else if (d.mode == 1u) acc[p] += nmda ? (c + d.boost) : c;
                        else acc[p] += nmda ? 1u : 0u;
                    }
                }
            }
        }
    }
    const uint32_t gid = base + n;
    const uint32_t bi = d.bias[gid] >> d.biasShift;
    uint32_t ub[PT], fs[PT];
#pragma unroll
    for (int p = 0; p < PT; ++p) {
        const uint32_t sl = min(slot0 + (uint32_t)p, d.nslots - 1u);
        ub[p] = d.st_bias[sl];
        fs[p] = d.facil != 0u ? d.spk[(size_t)sl * d.ncol + col] : 0u;
    }
#pragma unroll
    for (int p = 0; p < PT; ++p) {
        if (((live >> p) & 1u) == 0u) continue;
        const uint32_t sl = slot0 + (uint32_t)p;
        uint32_t a = acc[p];
        if (((fs[p] >> n) & 1u) != 0u) a += d.facil;
        if (ub[p] != 0u) a = (a > bi) ? (a - bi) : 0u;
        const size_t ai = (size_t)sl * d.ncol * NPC + gid;
        if (!hi) {
            if (d.acc16) reinterpret_cast<uint16_t*>(d.acc)[ai] = (uint16_t)a;
            else d.acc[ai] = a;
        }
        const uint32_t drv = half_sum(a);
        if (lane == 0u) d.drive[(size_t)sl * d.ncol + col] = drv;
    }
}

// k_probe_select: one workgroup per slot.  Column top-K theta, then the gated
// k-of-32 for columns that pass (the others are silent), the input clamp, the
// spike and history writes and the slot's group-activity bits.  The margin,
// stats and drive-log side outputs of k_emit are not produced for probes.
// `parts` 1024-thread workgroups per probe, each over its own range of the
// columns (a multiple of 64): the top-K over the probe's drives (every part
// computes it), then 16 waves over the part's columns, in groups of up to 4
// chunks of 64 (wave w takes chunks w, w + 16, ...).  A group's drive and
// clamp words are requested before they are needed (the first group's before
// the top-K), and the k-of-32 selection of the group's gated columns is
// batched over all its chunks: up to 8 gated columns per round, two at a time
// (half-wave h takes every other one), their acc loads all in flight
// together, results handed to the column's lane through LDS.  Split parts
// shorten a lightly loaded step; the last part to finish (a per-slot counter)
// advances the slot's step, after every part has read it.
constexpr uint32_t SEL_CHUNKS = 4;
struct SelGroup { uint32_t dv[SEL_CHUNKS], st[SEL_CHUNKS]; };
__device__ __forceinline__ SelGroup select_inputs(const ProbeDev& d, size_t so, uint32_t g0, uint32_t cEnd,
                                                  bool clampPhase) {
    SelGroup in;
#pragma unroll
    for (uint32_t j = 0; j < SEL_CHUNKS; ++j) {
        const uint32_t c = g0 + j * 1024u + lane_id();
        const bool inb = c < cEnd;
        in.dv[j] = inb ? d.drive[so + c] : 0u;
        in.st[j] = (inb && clampPhase && c < d.incols) ? d.stim[so + c] : 0u;
    }
    return in;
}
__global__ __launch_bounds__(1024) void k_probe_select(ProbeDev d, uint32_t parts) {
    __shared__ uint32_t bres[16][SEL_CHUNKS * 64];
    const uint32_t sl = blockIdx.x / parts, part = blockIdx.x % parts;
    if (sl >= d.nslots) return;
    const uint32_t t = d.st_t[sl];
    if (t >= d.st_T[sl]) return;
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = threadIdx.x >> 6, h = lane >> 5, n = lane & 31u;
    const size_t so = (size_t)sl * d.ncol;
    const bool clampPhase = t < d.st_clamp[sl];
    const uint32_t span = ((d.ncol + parts - 1u) / parts + 63u) & ~63u;
    const uint32_t cEnd = min(d.ncol, (part + 1u) * span);
    const uint32_t gFirst = part * span + wave * 64u;
    SelGroup in = select_inputs(d, so, gFirst, cEnd, clampPhase);  // in flight during the top-K
    const uint32_t th = topk_block<1024>(d.drive + so, part == 0u ? d.theta + 2u * sl : nullptr, d.ncol, d.incols,
                                         d.target, d.topkpar);
    const uint32_t* __restrict__ acc = d.acc + so * NPC;
    const uint16_t* __restrict__ acc16 = reinterpret_cast<const uint16_t*>(d.acc) + so * NPC;
    uint32_t* __restrict__ hist = d.hist + ((size_t)sl * d.maxt + t) * d.ncol;
    uint32_t* __restrict__ bw = bres[wave];
    for (uint32_t g0 = gFirst; g0 < cEnd; g0 += SEL_CHUNKS * 1024u) {
        if (g0 != gFirst) in = select_inputs(d, so, g0, cEnd, clampPhase);
        uint64_t m[SEL_CHUNKS];
#pragma unroll
        for (uint32_t j = 0; j < SEL_CHUNKS; ++j) {
            const uint32_t c = g0 + j * 1024u + lane;
            const bool clampCol = c < cEnd && clampPhase && c < d.incols;
            m[j] = __ballot(c < cEnd && !clampCol && in.dv[j] >= th);
        }
        while ((m[0] | m[1] | m[2] | m[3]) != 0ull) {  // wave-uniform
            // up to 8 gated columns (chunk * 64 + column, 256 = none), half-wave h taking every other
            // one; the picks are taken with constant indices only, so masks and picks stay in registers
            uint32_t ci[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                uint32_t pa = 256u, pb = 256u;
#pragma unroll
                for (uint32_t j = 0; j < SEL_CHUNKS; ++j) {
                    if (pa == 256u && m[j] != 0ull) {
                        pa = j * 64u + (uint32_t)__ffsll((unsigned long long)m[j]) - 1u;
                        m[j] &= m[j] - 1ull;
                    }
                }
#pragma unroll
                for (uint32_t j = 0; j < SEL_CHUNKS; ++j) {
                    if (pb == 256u && m[j] != 0ull) {
                        pb = j * 64u + (uint32_t)__ffsll((unsigned long long)m[j]) - 1u;
                        m[j] &= m[j] - 1ull;
                    }
                }
                ci[q] = h ? pb : pa;
            }
            uint32_t v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const size_t ai = (size_t)(g0 + (ci[q] >> 6) * 1024u + (ci[q] & 63u)) * NPC + n;
                v[q] = ci[q] < 256u ? (d.acc16 ? (uint32_t)acc16[ai] : acc[ai]) : 0u;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const Kwta kw = kwta32(v[q], d.k, n, h);
                if (n == 0u && ci[q] < 256u) bw[ci[q]] = kw.ball;
            }
        }
        uint32_t* ga = d.gact + (size_t)sl * d.gw;
#pragma unroll
        for (uint32_t j = 0; j < SEL_CHUNKS; ++j) {
            const uint32_t c0 = g0 + j * 1024u;
            if (c0 >= cEnd) break;  // wave-uniform
            const uint32_t c = c0 + lane;
            const bool inb = c < cEnd;
            const bool clampCol = inb && clampPhase && c < d.incols;
            uint32_t b = in.st[j];
            if (inb && !clampCol && in.dv[j] >= th) b = bw[j * 64u + lane];  // written above by this wave
            if (inb) { d.spk[so + c] = b; hist[c] = b; }
            const uint64_t nz = __ballot(inb && b != 0u);  // one bit per column of these 64
            if (lane == 0u) ga[c0 >> 5] = (uint32_t)nz;
            if (lane == 1u && c0 + 32u < d.ncol) ga[(c0 >> 5) + 1u] = (uint32_t)(nz >> 32);
        }
    }
    __syncthreads();
    if (tid == 0u) {
        if (parts == 1u) {
            d.st_t[sl] = t + 1u;
        } else if (atomicAdd(&d.selDone[sl], 1u) == parts - 1u) {
            d.selDone[sl] = 0u;
            d.st_t[sl] = t + 1u;
        }
    }
}

// Start of a probe: zero spikes and spiking-column bits, step counter to 0.
__global__ __launch_bounds__(256) void k_probe_reset(ProbeDev d, uint32_t first, uint32_t count) {
    const uint32_t sl = first + blockIdx.x;
    if (blockIdx.x >= count || sl >= d.nslots) return;
    for (uint32_t i = threadIdx.x; i < d.ncol; i += 256u) d.spk[(size_t)sl * d.ncol + i] = 0u;
    for (uint32_t i = threadIdx.x; i < d.gw; i += 256u) d.gact[(size_t)sl * d.gw + i] = 0u;
    if (threadIdx.x == 0u) d.st_t[sl] = 0u;
}

// ============================================================== many networks
// Lock-step training of independent networks: one launch covers every
// network that still has work.  The per-network work queues are a device
// table sched[round][net] (pattern index, or ~0 once the network is done) and
// a device round counter that the last kernel of a presentation advances, so
// one captured presentation graph is replayed for every round without host
// writes in between.
__device__ __forceinline__ uint32_t mb_pattern(const MultiDev& m, uint32_t net) {
    return m.sched[(size_t)(*m.round) * m.nnet + net];
}

// The descriptor of network `net` with its `active` pattern: all fields are
// loaded in one batch before the idle test (the asm makes the kernel wait for
// them there), instead of the test's load and then the descriptor's.
__device__ __forceinline__ NetDev mb_net(const MultiDev& m, uint32_t net) {
    const NetDev d = m.nets[net];
    asm volatile("" ::"s"(d.W0), "s"(d.W3), "s"(d.conn), "s"(d.spk[0]), "s"(d.drive), "s"(d.cntProp),
                 "s"(d.cntLearn), "s"(d.ncol), "s"(d.active));
    return d;
}

__global__ __launch_bounds__(256) void k_mb_begin(MultiDev m) {
    const uint32_t net = blockIdx.y;
    const uint32_t p = mb_pattern(m, net);
    if (blockIdx.x == 0u && threadIdx.x == 0u) const_cast<NetDev*>(m.nets)[net].active = p;
    if (p == 0xFFFFFFFFu) return;
    const NetDev d = m.nets[net];
    const uint32_t* tab = reinterpret_cast<const uint32_t*>(m.stimTabs[net]) + (size_t)p * d.ncol;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < d.ncol) { d.stim[i] = tab[i]; d.spk[0][i] = 0u; d.spk[1][i] = 0u; d.colNZ[i] = 0u; }
    if (i < d.nstats) d.stats[i] = 0u;
}

// Columns per batched propagation workgroup (two per wave).  Measured on 25
// networks: 8 columns 906 ms, 16 columns 955-966 ms, 32 columns 1073-1083 ms
// (with the emits' columns on the propagation's XCD or not).
constexpr uint32_t MB_DRIVE_COLS = 8;

// First column of wave w of a batched emit workgroup b (8 columns, two per
// wave): the emit workgroups on XCD b % 8 take the columns that the
// propagation workgroups on that XCD ran (MB_DRIVE_COLS columns each, split
// into parts of 8).
__device__ __forceinline__ uint32_t mb_emit_col(uint32_t b, uint32_t w) {
    constexpr uint32_t H = MB_DRIVE_COLS / 8u;  // emit workgroups per propagation workgroup
    const uint32_t r = b >> 3, j = (b & 7u) + 8u * (r / H);
    return MB_DRIVE_COLS * j + (r % H) * 8u + 2u * w;
}
union MbDriveLds {
    VoteList<MB_DRIVE_COLS> vl;
    uint32_t items[MB_DRIVE_COLS / 2][2 * ITEMS_PER_COL];
};
static_assert(offsetof(VoteList<MB_DRIVE_COLS>, nt) >= sizeof(uint32_t) * (MB_DRIVE_COLS / 2) * 2u * ITEMS_PER_COL,
              "the vote counts must not share LDS with the item lists");

// Batched propagation, MB_DRIVE_COLS columns per workgroup (two per wave).  FUSE: the
// deferred plasticity of timestep t - 1 (parameters PL) first rewrites the
// voted tiles of the workgroup's columns, as in k_drive.  FIRST: step 1 of a
// round whose step 0 (parameters PL) is clamped: as k_drive<MODE, false,
// true>, with the round's work of k_mb_emit_first (the pattern from the
// schedule into the descriptor, its clamp row into the stimulus), whose
// clamp row is step 0's spike vector.
template <int MODE, bool FUSE, bool FIRST = false>
__global__ __launch_bounds__(32 * MB_DRIVE_COLS) void k_mb_drive(MultiDev m, Params P, uint32_t t, Params PL) {
    const uint32_t net = blockIdx.y;
    uint32_t p = 0u;
    if (FIRST) {
        p = mb_pattern(m, net);
        if (blockIdx.x == 0u && threadIdx.x == 0u) const_cast<NetDev*>(m.nets)[net].active = p;
        if (p == 0xFFFFFFFFu) return;  // workgroup-uniform
    }
    const NetDev d = FIRST ? m.nets[net] : mb_net(m, net);  // FIRST: its `active` may be the last round's, unread
    if (!FIRST && d.active == 0xFFFFFFFFu) return;  // workgroup-uniform
    extern __shared__ uint4 fired[];  // fired_lds_bytes(ncol)
    // The vote list is used before the propagation's item lists: they share
    // LDS (8 KiB less per workgroup, so LDS does not bound the occupancy below
    // the registers; 25 networks 960 -> 942 ms), all but the counts, which
    // every wave reads after the vote whether or not it rewrites.
    __shared__ MbDriveLds sh;
    VoteList<MB_DRIVE_COLS>& vl = sh.vl;
    const uint32_t wave = threadIdx.x >> 6, lane = lane_id(), h = lane >> 5;
    const uint32_t colA = blockIdx.x * MB_DRIVE_COLS + wave * 2u;
    const uint32_t col = colA + h, ccol = min(col, d.ncol - 1u);
    uint32_t* const s0 = d.spk[0];  // no dynamic index into the local copy
    uint32_t* const s1 = d.spk[1];
    const uint32_t* tab = FIRST ? reinterpret_cast<const uint32_t*>(m.stimTabs[net]) + (size_t)p * d.ncol : nullptr;
    const uint32_t* in = FIRST ? tab : (t & 1u) ? s1 : s0;
    DriveIn di = drive_inputs(d, in, ccol);  // di.prev: the column's spike word of step t - 1
    VoteIn vi{};
    if (FUSE) vi = vote_inputs(d, ccol, reinterpret_cast<const uint2*>(d.cntLearn));
    if (FUSE && threadIdx.x == 0u) { vl.nt = 0u; vl.nl = 0u; }
    if (FIRST) {  // step 0's work for this half's column (k_mb_emit_first)
        const uint32_t n = lane & 31u;
        if (PL.learn != 0u && ccol >= PL.inCols) {
            const int dec = ((PL.step % max(PL.biasDownEvery, 1u)) == 0u) ? (int)PL.biasDown : 0;
            di.biasw = (uint32_t)max(0, (int)di.biasw - dec);
            if (col < d.ncol) d.bias[col * NPC + n] = di.biasw;
        }
        if (col < d.ncol && n == 0u) {
            d.stim[col] = di.prev;  // read by the clamped steps that follow
            s1[col] = di.prev;
            if (PL.histOn != 0u) d.hist[(size_t)PL.step * d.ncol + col] = di.prev;
        }
    }
    const uint8_t* f = stage_fired(d, fired, FUSE, tab, PL.inCols);
    if (FUSE) {
        if (col < d.ncol && learns(PL, di.prev, vi.marg))
            vote_column<MB_DRIVE_COLS>(d, PL, col, di.prev, di.gl, vi, h, reinterpret_cast<uint2*>(d.cntLearn), vl);
        __syncthreads();
        if (vl.nl != 0u) {  // workgroup-uniform
            // two lines per half-wave and round: 55 registers instead of 85 with four, and a many-network
            // launch is bound by how many of its workgroups are resident (measured: 25 networks 968 ms
            // against 1006 ms; eight lines, 145 registers: 1261 ms)
            rewrite_voted<MB_DRIVE_COLS, 2>(d, PL, (t & 1u) ? s0 : s1, vl, wave, MB_DRIVE_COLS / 2u,
                                            blockIdx.x * MB_DRIVE_COLS * NSLOT);
            __syncthreads();  // the rewritten lines are read below, by any wave of the workgroup
        }
    }
    if (colA >= d.ncol) return;
    drive_pair<MODE>(d, P, in, colA, di, f, sh.items[wave]);
}

// 256 threads per network: this launch is a latency chain of a few barriers,
// cheaper with 4 waves than with 16 or 2, and it leaves the device to the
// other stream groups' kernels.  Measured on 25 networks: 256 threads 906 ms,
// 1024 threads 947 ms, 128 threads 918 ms.  Without this launch, the emit
// gating each column by counting the drives above it (gate_by_count, as for
// one network) took 1150 ms: each 8-column workgroup then reads its
// network's whole drive vector.
__global__ __launch_bounds__(256) void k_mb_topk(MultiDev m, Params P) {
    const uint32_t net = blockIdx.x;
    const NetDev d = mb_net(m, net);
    if (d.active == 0xFFFFFFFFu) return;
    topk_block<256>(d.drive, d.theta, d.ncol, P.inCols, P.targetCols, d.topkpar);
}

// k_emit + k_bias (as emit_cols) for two columns per wave, the form for many
// networks at once: half h is column colA + h, lane n of a half is neuron n.
// No side statistics.  Returns the half's column's spike word and margin flag.
template <bool LEARN>
__device__ __forceinline__ uint2 emit_pair(const NetDev& d, const Params& P, uint32_t* __restrict__ spkOut,
                                           uint32_t colA) {
    const uint32_t lane = lane_id(), h = lane >> 5, n = lane & 31u;
    const uint32_t col = colA + h;
    const uint32_t gid = col * NPC + n;
    const bool clampCol = P.clamp != 0u && col < P.inCols;
    const bool rec = col >= P.inCols;
    const uint32_t dv = d.drive[col];
    const uint32_t v = d.acc16 ? (uint32_t)reinterpret_cast<const uint16_t*>(d.acc)[gid] : d.acc[gid];
    const uint32_t stimW = clampCol ? d.stim[col] : 0u;
    const uint32_t bia = (LEARN && rec) ? d.bias[gid] : 0u;
    const bool gatePass = dv >= max(d.theta[0], 1u);
    uint32_t b = 0u, vk = 0u, vk1 = 0u;
    if (clampCol || gatePass) {
        const Kwta kw = kwta32(v, P.k, n, h);
        vk = kw.vk; vk1 = kw.vk1;
        b = gatePass ? kw.ball : 0u;
        if (clampCol) b = stimW;
    }
    const bool post = ((b >> n) & 1u) != 0u;
    if (n == 0u) {
        spkOut[col] = b;
        d.marg[col] = (vk > vk1) ? 1u : 0u;
        d.colNZ[col] = b != 0u ? 1u : 0u;
        if (P.histOn != 0u) d.hist[(size_t)P.step * d.ncol + col] = b;
    }
    if (LEARN && rec) {  // k_bias
        const int dec = ((P.step % max(P.biasDownEvery, 1u)) == 0u) ? (int)P.biasDown : 0;
        const int nb = (int)bia + (post ? (int)P.biasUp : -dec);
        d.bias[gid] = (uint32_t)max(0, nb);
    }
    // this step's plasticity runs at the next propagation (k_mb_drive) or, at the round's last step,
    // right after the emit (k_mb_emit_learn)
    return make_uint2(b, (vk > vk1) ? 1u : 0u);
}

template <bool LEARN>
__global__ __launch_bounds__(256) void k_mb_emit(MultiDev m, Params P, uint32_t t) {
    const uint32_t net = blockIdx.y;
    const NetDev d = mb_net(m, net);
    if (d.active == 0xFFFFFFFFu) return;
    uint32_t* const s0 = d.spk[0];
    uint32_t* const s1 = d.spk[1];
    const uint32_t colA = mb_emit_col(blockIdx.x, threadIdx.x >> 6);  // 8 columns per workgroup
    if (colA >= d.ncol) return;
    (void)emit_pair<LEARN>(d, P, (t & 1u) ? s0 : s1, colA);
}

// The last step t of a round, when it learns (as k_emit_learn for one
// network): k_mb_emit, then the step's plasticity of the workgroup's 8
// columns in the same workgroup, two lines per half-wave and round (as
// k_mb_drive).  The round's last launch: it advances the round counter
// (read by the next round's first launch only).  Its work is counted in
// cntLast.  The rewrite's other forms (4 lines per half-wave and round,
// issued round by round or pipelined) measured the same on 25 networks
// (901-911 ms).
__global__ __launch_bounds__(256) void k_mb_emit_learn(MultiDev m, Params P, uint32_t t) {
    if (blockIdx.x == 0u && blockIdx.y == 0u && threadIdx.x == 0u) *m.round += 1u;
    const uint32_t net = blockIdx.y;
    const NetDev d = mb_net(m, net);
    if (d.active == 0xFFFFFFFFu) return;  // workgroup-uniform
    __shared__ VoteList<8> vl;
    uint32_t* const s0 = d.spk[0];
    uint32_t* const s1 = d.spk[1];
    const uint32_t wave = threadIdx.x >> 6, lane = lane_id(), h = lane >> 5;
    const uint32_t colA = mb_emit_col(blockIdx.x, wave);  // 8 columns per workgroup
    const uint32_t col = colA + h, ccol = min(col, d.ncol - 1u);
    const uint32_t gl = d.conn[ccol * NSLOT + (lane & 31u)];
    const VoteIn vi = vote_inputs(d, ccol, reinterpret_cast<const uint2*>(d.cntLast));
    if (threadIdx.x == 0u) { vl.nt = 0u; vl.nl = 0u; }
    __syncthreads();
    const uint2 bm = colA < d.ncol ? emit_pair<true>(d, P, (t & 1u) ? s0 : s1, colA) : make_uint2(0u, 0u);
    if (col < d.ncol && learns(P, bm.x, bm.y))
        vote_column<8>(d, P, col, bm.x, gl, vi, h, reinterpret_cast<uint2*>(d.cntLast), vl);
    __syncthreads();
    if (vl.nl != 0u)  // workgroup-uniform
        rewrite_voted<8, 2>(d, P, (t & 1u) ? s1 : s0, vl, wave, 4u, mb_emit_col(blockIdx.x, 0u) * NSLOT);
}

// The first timestep of a presentation for many networks (the batched form
// of k_emit_first, which the network's own begin work joins): this round's
// pattern is read from the schedule and stored in the network's descriptor
// for the launches that follow, the pattern's clamp row becomes the
// network's stimulus, the plasticity list counts are cleared, and the step
// is emitted directly: only clamped input columns spike, every accumulator
// and drive being 0.
template <bool LEARN>
__global__ __launch_bounds__(256) void k_mb_emit_first(MultiDev m, Params P) {
    const uint32_t net = blockIdx.y;
    const uint32_t p = mb_pattern(m, net);
    if (blockIdx.x == 0u && threadIdx.x == 0u) const_cast<NetDev*>(m.nets)[net].active = p;
    if (p == 0xFFFFFFFFu) return;
    const NetDev d = m.nets[net];  // its `active` may be the last round's; not read here
    const uint32_t lane = lane_id(), h = lane >> 5, n = lane & 31u;
    const uint32_t col = mb_emit_col(blockIdx.x, threadIdx.x >> 6) + h;  // 8 columns per workgroup
    if (col >= d.ncol) return;
    const uint32_t sw = reinterpret_cast<const uint32_t*>(m.stimTabs[net])[(size_t)p * d.ncol + col];
    const uint32_t b = (P.clamp != 0u && col < P.inCols) ? sw : 0u;
    if (n == 0u) {
        d.stim[col] = sw;  // read by the clamped steps that follow
        d.spk[1][col] = b;
        d.marg[col] = 0u;
        d.colNZ[col] = b != 0u ? 1u : 0u;
        if (P.histOn != 0u) d.hist[(size_t)P.step * d.ncol + col] = b;
    }
    if (LEARN && col >= P.inCols) {  // k_bias (recurrent columns never spike here)
        const uint32_t gid = col * NPC + n;
        const int dec = ((P.step % max(P.biasDownEvery, 1u)) == 0u) ? (int)P.biasDown : 0;
        d.bias[gid] = (uint32_t)max(0, (int)d.bias[gid] - dec);
    }
}

__global__ void k_mb_end(MultiDev m) {
    if (threadIdx.x == 0u) *m.round += 1u;
}

// Order-independent checksum: sum of w_i * (2i + 1) + hash32(w_i ^ i) (mod 2^64).
__global__ __launch_bounds__(256) void k_checksum(const uint32_t* __restrict__ p, size_t n,
                                                  unsigned long long* out) {
    unsigned long long acc = 0;
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256ull) {
        const uint32_t w = p[i];
        acc += (unsigned long long)w * (2ull * i + 1ull) + hash32(w ^ (uint32_t)i ^ (uint32_t)(i >> 32));
    }
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane_id() == 0u) atomicAdd(out, acc);
}

// Serving: scatter a packed list of admitted requests into their slots.
// Entry layout: [slot, T, clamp, useBias, stim[ncol]].
__global__ __launch_bounds__(256) void k_probe_admit(ProbeDev d, const uint32_t* __restrict__ list, uint32_t n) {
    if (blockIdx.x >= n) return;
    const uint32_t* e = list + (size_t)blockIdx.x * (4u + d.ncol);
    const uint32_t sl = e[0];
    if (sl >= d.nslots) return;
    for (uint32_t i = threadIdx.x; i < d.ncol; i += 256u) {
        d.stim[(size_t)sl * d.ncol + i] = e[4 + i];
        d.spk[(size_t)sl * d.ncol + i] = 0u;
    }
    for (uint32_t i = threadIdx.x; i < d.gw; i += 256u) d.gact[(size_t)sl * d.gw + i] = 0u;
    if (threadIdx.x == 0u) {
        d.st_T[sl] = e[1]; d.st_clamp[sl] = e[2]; d.st_bias[sl] = e[3]; d.st_t[sl] = 0u;
    }
}

// Serving: pack finished slots' results: final spikes, then one 64-bit
// checksum per history row (the k_checksum formula over the row's words).
// out + offs[i] receives ncol words followed by 2*T words.
// Workgroup (j, y): slot slots[j]; y == 0 also copies its spikes; row y of
// its history (y < T) is reduced to its 64-bit digest.
__global__ __launch_bounds__(256) void k_probe_gather(ProbeDev d, const uint32_t* __restrict__ slots,
                                                      const uint32_t* __restrict__ offs, uint32_t n,
                                                      uint32_t* __restrict__ out) {
    __shared__ unsigned long long part[4];
    if (blockIdx.x >= n) return;
    const uint32_t sl = slots[blockIdx.x], t = blockIdx.y;
    uint32_t* o = out + offs[blockIdx.x];
    if (t == 0u)
        for (uint32_t i = threadIdx.x; i < d.ncol; i += 256u) o[i] = d.spk[(size_t)sl * d.ncol + i];
    if (t >= d.st_T[sl]) return;  // uniform over the workgroup
    const uint32_t* row = d.hist + ((size_t)sl * d.maxt + t) * d.ncol;
    unsigned long long acc = 0;
    for (uint32_t i = threadIdx.x; i < d.ncol; i += 256u) {
        const uint32_t w = row[i];
        acc += (unsigned long long)w * (2ull * i + 1ull) + hash32(w ^ i);
    }
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane_id() == 0u) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0u) {
        const unsigned long long v = part[0] + part[1] + part[2] + part[3];
        o[d.ncol + 2u * t] = (uint32_t)v;
        o[d.ncol + 2u * t + 1u] = (uint32_t)(v >> 32);
    }
}

inline uint32_t blocks_for(size_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

}  // namespace

// ============================================================== launchers
hipError_t launch_init(const NetDev& d, uint32_t nand, uint32_t seed, hipStream_t s) {
    const uint32_t planeU4 = d.ncol * NSLOT * (TILE_WORDS / 4u);  // one thread per reference uint4
    k_init_w<<<blocks_for(planeU4, 256), 256, 0, s>>>(d.W0, d.W1, d.W2, d.W3, nand, seed, planeU4);
    const uint32_t nconn = d.ncol * NSLOT;
    k_init_conn<<<blocks_for(nconn, 256), 256, 0, s>>>(d.conn, d.ncol / COLS_PER_GRP, d.incols / COLS_PER_GRP,
                                                        d.inbr, seed, nconn);
    return hipGetLastError();
}

hipError_t launch_fill(uint32_t* p, uint32_t value, size_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    const uint32_t blocks = (uint32_t)std::min<size_t>(blocks_for(n, 256), 4096);
    k_fill<<<blocks, 256, 0, s>>>(p, value, n);
    return hipGetLastError();
}

hipError_t launch_clear(const NetDev& d, hipStream_t s) {
    const uint32_t n = std::max(d.ncol, d.nstats);
    k_clear<<<blocks_for(n, 256), 256, 0, s>>>(d.spk[0], d.spk[1], d.ncol, d.stats, d.nstats, d.colNZ, d.learnCnt);
    return hipGetLastError();
}

// With timing events, the launch carries them in its dispatch packet
// (hipExtLaunchKernel), so they bracket the kernel itself and not the queue
// traffic around it.
template <class F, class... A>
static hipError_t launch(F k, dim3 g, dim3 b, uint32_t lds, hipStream_t s, hipEvent_t e0, hipEvent_t e1, A... a) {
    if (e0) hipExtLaunchKernelGGL(k, g, b, lds, s, e0, e1, 0, a...);
    else hipLaunchKernelGGL(k, g, b, lds, s, a...);
    return hipGetLastError();
}

hipError_t launch_drive(const NetDev& d, const Params& P, uint32_t t, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                        const Params* PL, const Params* P0, const uint32_t* stim0) {
    const uint32_t* in = d.spk[t & 1u];
    const uint32_t* prevIn = d.spk[(t & 1u) ^ 1u];  // step t - 1's presynaptic spikes (its plasticity reads them)
    const dim3 g(blocks_for(d.ncol, DRIVE_COLS));
    const dim3 bl(64u * DRIVE_COLS);
    const uint32_t lds = fired_lds_bytes(d.ncol);
    if (P0) {  // step 1 taking step 0's work (k_drive<MODE, false, true>)
        const uint32_t* st = stim0 ? stim0 : d.stim;
        if (P.mode == 0u) return launch(k_drive<0, false, true>, g, bl, lds, s, e0, e1, d, P, st, *P0, in);
        if (P.mode == 1u) return launch(k_drive<1, false, true>, g, bl, lds, s, e0, e1, d, P, st, *P0, in);
        return launch(k_drive<2, false, true>, g, bl, lds, s, e0, e1, d, P, st, *P0, in);
    }
    if (PL) {
        if (P.mode == 0u) return launch(k_drive<0, true>, g, bl, lds, s, e0, e1, d, P, in, *PL, prevIn);
        if (P.mode == 1u) return launch(k_drive<1, true>, g, bl, lds, s, e0, e1, d, P, in, *PL, prevIn);
        return launch(k_drive<2, true>, g, bl, lds, s, e0, e1, d, P, in, *PL, prevIn);
    }
    if (P.mode == 0u) return launch(k_drive<0, false>, g, bl, lds, s, e0, e1, d, P, in, P, prevIn);
    if (P.mode == 1u) return launch(k_drive<1, false>, g, bl, lds, s, e0, e1, d, P, in, P, prevIn);
    return launch(k_drive<2, false>, g, bl, lds, s, e0, e1, d, P, in, P, prevIn);
}

hipError_t launch_learn(const NetDev& d, const Params& P, uint32_t t, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    const uint32_t* in = d.spk[t & 1u];
    const uint32_t* out = d.spk[(t & 1u) ^ 1u];
    const dim3 g(std::min<uint32_t>(d.ncol, 512u));  // grid-stride over the listed columns
    return launch(k_learn<true>, g, dim3(256), 0u, s, e0, e1, d, P, in, out);
}

hipError_t launch_emit_learn(const NetDev& d, const Params& P, uint32_t t, hipStream_t s, hipEvent_t e0,
                             hipEvent_t e1) {
    constexpr uint32_t W = EMIT_LEARN_WAVES;
    const dim3 g(blocks_for(d.ncol, 8u * W) * 8u), b(64u * W);  // as launch_emit
    return launch(k_emit_learn<W>, g, b, 0u, s, e0, e1, d, P, d.spk[t & 1u], d.spk[(t & 1u) ^ 1u]);
}

hipError_t launch_emit(const NetDev& d, const Params& P, uint32_t t, bool learn, bool stats, hipStream_t s,
                       hipEvent_t e0, hipEvent_t e1, bool first) {
    const uint32_t* in = d.spk[t & 1u];
    uint32_t* out = d.spk[(t & 1u) ^ 1u];
    const dim3 g(blocks_for(d.ncol, 128) * 8u), b(1024);  // 16 columns per workgroup, XCD-aligned (emit_col)
    if (first) {  // step 0 without side statistics (k_emit_first)
        if (learn) return launch(k_emit_first<true>, g, b, 0u, s, e0, e1, d, P, out);
        return launch(k_emit_first<false>, g, b, 0u, s, e0, e1, d, P, out);
    }
    if (learn && stats) return launch(k_emit<true, true>, g, b, 0u, s, e0, e1, d, P, in, out);
    if (learn) {
        constexpr uint32_t W = EMIT_TRAIN_WAVES;
        return launch(k_emit<true, false, W>, dim3(blocks_for(d.ncol, 8u * W) * 8u), dim3(64u * W), 0u, s, e0, e1,
                      d, P, in, out);
    }
    if (stats) return launch(k_emit<false, true>, g, b, 0u, s, e0, e1, d, P, in, out);
    return launch(k_emit<false, false>, g, b, 0u, s, e0, e1, d, P, in, out);
}

hipError_t launch_wstat(const NetDev& d, uint32_t* out, hipStream_t s) {
    const size_t n = (size_t)d.ncol * NSLOT * (TILE_WORDS / 4u);
    k_wstat<<<2048, 256, 0, s>>>(reinterpret_cast<const uint4*>(d.W3), n, out);
    return hipGetLastError();
}

hipError_t launch_within(const NetDev& d, const uint32_t* ball, const uint32_t* postMask, uint32_t useMask,
                         uint32_t* out, hipStream_t s) {
    k_within<<<blocks_for(d.ncol, 4), 256, 0, s>>>(d, ball, postMask, useMask, out);  // a wave per column
    return hipGetLastError();
}

hipError_t launch_indeg(const NetDev& d, uint32_t* out, hipStream_t s) {
    k_neuron_counts<0><<<blocks_for(d.ncol, 4), 256, 0, s>>>(d, out);
    return hipGetLastError();
}

hipError_t launch_untouched(const NetDev& d, uint32_t* out, hipStream_t s) {
    k_neuron_counts<1><<<blocks_for(d.ncol, 4), 256, 0, s>>>(d, out);
    return hipGetLastError();
}

hipError_t launch_wdiag(const NetDev& d, const uint32_t* memb, const uint32_t* asmSpk, uint32_t nstim,
                        uint32_t sel, uint32_t* anyAsm, uint32_t* hist, hipStream_t s) {
    k_asm_union<<<blocks_for(d.ncol, 256), 256, 0, s>>>(asmSpk, nstim, d.ncol, anyAsm);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_wdiag<<<blocks_for(d.ncol, 4), 256, 0, s>>>(d, memb, asmSpk, anyAsm, nstim, sel, hist);
    return hipGetLastError();
}

hipError_t launch_bstest(uint32_t op, uint32_t lo, uint32_t hi, uint32_t* a0, uint32_t* a1, uint32_t* a2,
                         uint32_t* a3, const uint32_t* m, size_t n, hipStream_t s) {
    k_bstest<<<blocks_for(n, 256), 256, 0, s>>>(op, lo, hi, a0, a1, a2, a3, m, n);
    return hipGetLastError();
}

hipError_t launch_kwta_test(const uint32_t* v, uint32_t ncols, uint32_t k, uint32_t* ball, uint32_t* vk,
                            uint32_t* vk1, hipStream_t s) {
    k_kwta_test<<<blocks_for(ncols, 4), 256, 0, s>>>(v, ncols, k, ball, vk, vk1);
    return hipGetLastError();
}

hipError_t launch_topk_test(const uint32_t* drive, uint32_t ninst, uint32_t ncol, uint32_t incols, uint32_t target,
                            uint32_t par, uint32_t* theta, hipStream_t s) {
    k_topk_test<<<ninst, 1024, 0, s>>>(drive, theta, ncol, incols, target, par);
    return hipGetLastError();
}

}  // namespace axon

namespace axon {

// Dynamic LDS that limits a kernel to wgPerCU workgroups per CU (0: no limit):
// each workgroup then holds 160 KiB / wgPerCU of the CU's LDS.
static uint32_t occupancy_pad(const void* kernel, uint32_t wgPerCU) {
    if (!wgPerCU) return 0u;
    hipFuncAttributes a{};
    if (hipFuncGetAttributes(&a, kernel) != hipSuccess) return 0u;
    const size_t want = (160u * 1024u) / wgPerCU;
    return want > a.sharedSizeBytes ? (uint32_t)(want - a.sharedSizeBytes) : 0u;
}

template <typename K, typename... A>
static void probe_launch(K kernel, uint32_t blocks, uint32_t wgPerCU, hipStream_t s, const ProbeDev& d, A... args) {
    const uint32_t pad = occupancy_pad(reinterpret_cast<const void*>(kernel), wgPerCU);
    kernel<<<blocks, 256, pad, s>>>(d, args...);
}

template <int PT>
static void probe_items_launch(uint32_t blocks, uint32_t wgPerCU, hipStream_t s, const ProbeDev& d, uint32_t ncg,
                               uint32_t ntiles, uint32_t tpw) {
    if (d.mode == 0u) probe_launch(k_probe_items<PT, 0>, blocks, wgPerCU, s, d, ncg, ntiles, tpw);
    else if (d.mode == 1u) probe_launch(k_probe_items<PT, 1>, blocks, wgPerCU, s, d, ncg, ntiles, tpw);
    else probe_launch(k_probe_items<PT, 2>, blocks, wgPerCU, s, d, ncg, ntiles, tpw);
}

#ifndef AXON_PROBE_TPW
#define AXON_PROBE_TPW 4  // most probe tiles one k_probe_items wave walks
#endif

hipError_t launch_probe_step(const ProbeDev& d, uint32_t tile, uint32_t nact, uint32_t wavesPerCU, hipStream_t s) {
    nact = std::min(nact, d.nslots);
    if (nact == 0u) return hipSuccess;
    const uint32_t ncg = (d.ncol + 3u) / 4u;
    const uint32_t ntiles = (nact + tile - 1u) / tile;
    const uint32_t blocks = ((ncg + 7u) / 8u) * 8u * ntiles;
    const uint32_t wg = wavesPerCU / 4u;  // 4 waves per workgroup
    // tiles per wave: up to AXON_PROBE_TPW while the launch keeps >= 16384 waves (64 per CU)
    const uint32_t tpw = std::max<uint32_t>(1u, std::min<uint32_t>(AXON_PROBE_TPW, (uint32_t)((uint64_t)ntiles * d.ncol / 16384u)));
    const uint32_t iblocks = ((ncg + 7u) / 8u) * 8u * ((ntiles + tpw - 1u) / tpw);
    if ((tile == 1u || tile == 2u || tile == 4u || tile == 8u) && (uint64_t)tile * d.ncol >= (1u << 20)) {
        // item words hold spike indices below 2^20, and the pad item's index 2^20 - 1 must lie past a
        // tile's spike words: wider batches take the per-branch kernel (same results)
        const uint32_t nt16 = (nact + 15u) / 16u;
        probe_launch(k_probe_drive<16>, ((ncg + 7u) / 8u) * 8u * nt16, wg, s, d, ncg, nt16);
        tile = 0u;
    }
    switch (tile) {
        case 0: break;
        case 1: probe_items_launch<1>(iblocks, wg, s, d, ncg, ntiles, tpw); break;
        case 2: probe_items_launch<2>(iblocks, wg, s, d, ncg, ntiles, tpw); break;
        case 4: probe_items_launch<4>(iblocks, wg, s, d, ncg, ntiles, tpw); break;
        case 8: probe_items_launch<8>(iblocks, wg, s, d, ncg, ntiles, tpw); break;
        case 16: probe_launch(k_probe_drive<16>, blocks, wg, s, d, ncg, ntiles); break;
        case 32: probe_launch(k_probe_drive<32>, blocks, wg, s, d, ncg, ntiles); break;
        case 64: probe_launch(k_probe_drive<64>, blocks, wg, s, d, ncg, ntiles); break;
        default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // up to 64 live slots: each probe's select split over 4 workgroups
    const uint32_t parts = nact <= 64u ? 4u : 1u;
    k_probe_select<<<nact * parts, 1024, 0, s>>>(d, parts);
    return hipGetLastError();
}

hipError_t launch_probe_reset(const ProbeDev& d, uint32_t first, uint32_t count, hipStream_t s) {
    if (!count) return hipSuccess;
    k_probe_reset<<<count, 256, 0, s>>>(d, first, count);
    return hipGetLastError();
}

}  // namespace axon

namespace axon {

hipError_t launch_multi_begin(const MultiDev& m, uint32_t ncol, uint32_t nstats, hipStream_t s) {
    dim3 g((std::max(ncol, nstats) + 255u) / 256u, m.nnet);
    k_mb_begin<<<g, 256, 0, s>>>(m);
    return hipGetLastError();
}
hipError_t launch_multi_drive(const MultiDev& m, const Params& P, uint32_t t, uint32_t ncol, hipStream_t s,
                              const Params* PL, const Params* P0) {
    const dim3 g((ncol + MB_DRIVE_COLS - 1u) / MB_DRIVE_COLS, m.nnet);
    constexpr uint32_t NT = 32u * MB_DRIVE_COLS;
    const uint32_t lds = fired_lds_bytes(ncol);
    if (P0) {
        if (P.mode == 0u) k_mb_drive<0, false, true><<<g, NT, lds, s>>>(m, P, t, *P0);
        else if (P.mode == 1u) k_mb_drive<1, false, true><<<g, NT, lds, s>>>(m, P, t, *P0);
        else k_mb_drive<2, false, true><<<g, NT, lds, s>>>(m, P, t, *P0);
        return hipGetLastError();
    }
    if (PL) {
        if (P.mode == 0u) k_mb_drive<0, true><<<g, NT, lds, s>>>(m, P, t, *PL);
        else if (P.mode == 1u) k_mb_drive<1, true><<<g, NT, lds, s>>>(m, P, t, *PL);
        else k_mb_drive<2, true><<<g, NT, lds, s>>>(m, P, t, *PL);
    } else {
        if (P.mode == 0u) k_mb_drive<0, false><<<g, NT, lds, s>>>(m, P, t, P);
        else if (P.mode == 1u) k_mb_drive<1, false><<<g, NT, lds, s>>>(m, P, t, P);
        else k_mb_drive<2, false><<<g, NT, lds, s>>>(m, P, t, P);
    }
    return hipGetLastError();
}
hipError_t launch_multi_topk(const MultiDev& m, const Params& P, hipStream_t s) {
    k_mb_topk<<<m.nnet, 256, 0, s>>>(m, P);
    return hipGetLastError();
}
hipError_t launch_multi_emit(const MultiDev& m, const Params& P, uint32_t t, bool learn, uint32_t ncol,
                             hipStream_t s) {
    if (learn) k_mb_emit<true><<<dim3((ncol + 7u) / 8u, m.nnet), 256, 0, s>>>(m, P, t);
    else k_mb_emit<false><<<dim3((ncol + 7u) / 8u, m.nnet), 256, 0, s>>>(m, P, t);
    return hipGetLastError();
}
hipError_t launch_multi_emit_learn(const MultiDev& m, const Params& P, uint32_t t, uint32_t ncol, hipStream_t s) {
    k_mb_emit_learn<<<dim3((ncol + 7u) / 8u, m.nnet), 256, 0, s>>>(m, P, t);
    return hipGetLastError();
}
hipError_t launch_multi_emit_first(const MultiDev& m, const Params& P, bool learn, uint32_t ncol, hipStream_t s) {
    if (learn) k_mb_emit_first<true><<<dim3((ncol + 7u) / 8u, m.nnet), 256, 0, s>>>(m, P);
    else k_mb_emit_first<false><<<dim3((ncol + 7u) / 8u, m.nnet), 256, 0, s>>>(m, P);
    return hipGetLastError();
}
hipError_t launch_multi_end(const MultiDev& m, hipStream_t s) {
    k_mb_end<<<1, 64, 0, s>>>(m);
    return hipGetLastError();
}

}  // namespace axon

// ============================================================== memory probes
// Read bandwidth of the access shapes the engine uses, to calibrate its
// percent-of-peak figures: a streaming pass (16 B per lane), 128-byte line
// gathers (a half-wave per line, the propagation's W3 reads) and 1 KiB tile
// gathers (a wave per tile), at pseudo-random unit-aligned places.  Each
// result is folded into a value that is stored only if it matches a constant
// that a fill of zeros never produces, so the loads stay and nothing is
// written.
namespace axon {
namespace {
__global__ __launch_bounds__(256) void k_mem_stream(const uint4* __restrict__ p, size_t n16, uint32_t* out) {
    uint32_t x = 0;
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n16; i += (size_t)gridDim.x * 256ull) {
        const uint4 v = p[i];
        x ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    if (x == 0x9E3779B9u) out[0] = x;
}
// unit = 128: lane (h, n) reads dword n of line (wave, it, h); unit = 1024: lane l
// reads 16 B at 16 l of tile (wave, it).  nunits = buffer bytes / unit.
template <uint32_t UNIT>
__global__ __launch_bounds__(256) void k_mem_gather(const uint32_t* __restrict__ p, uint32_t nunits, uint32_t iters,
                                                    uint32_t* out) {
    const uint32_t lane = lane_id(), wave = (blockIdx.x * 256u + threadIdx.x) >> 6;
    uint32_t x = 0;
#pragma unroll 8
    for (uint32_t it = 0; it < iters; ++it) {
        if (UNIT == 128u) {
            const uint32_t u = hash32(wave * 2654435761u + it * 40503u + (lane >> 5)) % nunits;
            x ^= p[(size_t)u * 32u + (lane & 31u)];
        } else {
            const uint32_t u = hash32(wave * 2654435761u + it * 40503u) % nunits;
            const uint4 v = reinterpret_cast<const uint4*>(p + (size_t)u * 256u)[lane];
            x ^= v.x ^ v.y ^ v.z ^ v.w;
        }
    }
    if (x == 0x9E3779B9u) out[0] = x;
}
}  // namespace
// bytes read by one launch
size_t mem_probe_bytes(uint32_t kind, size_t bytes, uint32_t blocks, uint32_t iters) {
    if (kind == 0) return bytes;
    return (size_t)blocks * 4u * iters * (kind == 1 ? 2u * 128u : 1024u);
}
// Cost of cross-workgroup synchronisation at a propagation launch's shape
// (the measurement behind running a timestep as separate launches rather
// than one persistent kernel with grid-wide waits): every workgroup of group
// y stores what a batched propagation workgroup does (8 column words and 256
// neuron words), then KIND 1 adds one arrival atomic per workgroup on the
// group's counter (what a grid barrier or a last-workgroup reduction needs),
// KIND 2 an agent-scope release fence before it (so that the stores are
// visible to the other XCDs' workgroups), KIND 3 a full agent-scope fence.
// Nothing waits for anything: every workgroup finishes on its own.
template <int KIND>
__global__ __launch_bounds__(256) void k_sync_probe(uint32_t* __restrict__ out, uint32_t* __restrict__ cnt,
                                                    uint32_t nwg) {
    const uint32_t g = blockIdx.y, b = blockIdx.x;
    uint32_t* o = out + ((size_t)g * nwg + b) * 264u;
    o[threadIdx.x] = threadIdx.x ^ b;
    if (threadIdx.x < 8u) o[256u + threadIdx.x] = b + threadIdx.x;
    if (KIND == 0) return;
    if (KIND == 2) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (KIND == 3) __threadfence();
    __syncthreads();
    if (threadIdx.x == 0u) atomicAdd(&cnt[g * 32u], 1u);
}

hipError_t launch_sync_probe(uint32_t kind, uint32_t* out, uint32_t* cnt, uint32_t groups, uint32_t nwg,
                             hipStream_t s) {
    const dim3 g(nwg, groups);
    if (kind == 0) k_sync_probe<0><<<g, 256, 0, s>>>(out, cnt, nwg);
    else if (kind == 1) k_sync_probe<1><<<g, 256, 0, s>>>(out, cnt, nwg);
    else if (kind == 2) k_sync_probe<2><<<g, 256, 0, s>>>(out, cnt, nwg);
    else k_sync_probe<3><<<g, 256, 0, s>>>(out, cnt, nwg);
    return hipGetLastError();
}
hipError_t launch_mem_probe(uint32_t kind, const uint32_t* p, size_t bytes, uint32_t blocks, uint32_t iters,
                            uint32_t* out, hipStream_t s) {
    if (kind == 0) k_mem_stream<<<blocks, 256, 0, s>>>(reinterpret_cast<const uint4*>(p), bytes / 16u, out);
    else if (kind == 1) k_mem_gather<128><<<blocks, 256, 0, s>>>(p, (uint32_t)(bytes / 128u), iters, out);
    else k_mem_gather<1024><<<blocks, 256, 0, s>>>(p, (uint32_t)(bytes / 1024u), iters, out);
    return hipGetLastError();
}
}  // namespace axon

namespace axon {
hipError_t launch_checksum(const uint32_t* p, size_t n, unsigned long long* out, hipStream_t s) {
    k_checksum<<<2048, 256, 0, s>>>(p, n, out);
    return hipGetLastError();
}
}  // namespace axon

namespace axon {
hipError_t launch_probe_admit(const ProbeDev& d, const uint32_t* list, uint32_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    k_probe_admit<<<n, 256, 0, s>>>(d, list, n);
    return hipGetLastError();
}
hipError_t launch_probe_gather(const ProbeDev& d, const uint32_t* slots, const uint32_t* offs, uint32_t n,
                               uint32_t* out, hipStream_t s) {
    if (!n) return hipSuccess;
    k_probe_gather<<<dim3(n, std::max(d.maxt, 1u)), 256, 0, s>>>(d, slots, offs, n, out);
    return hipGetLastError();
}
}  // namespace axon

namespace axon {

// ============================================================== allocation-free probe runs
// Probe state for callers that cannot allocate device memory: one static
// buffer per GPU, NOALLOC_SCRATCH_WORDS words, held until the process ends.
// Host code never passes `scratch` itself (its host-side name is not a device
// address); it passes the address noalloc_scratch() obtained from
// hipGetSymbolAddress on that device.
namespace {
__device__ __align__(256) uint32_t scratch[NOALLOC_SCRATCH_WORDS];

// Start of an allocation-free run: every slot's spikes and spiking-column
// bits to zero, its step counter, split-select counter and theta to zero,
// and its step count (T for slots below `count`, 0 = idle above), clamped
// steps and bias switch set.  One workgroup per slot.
__global__ __launch_bounds__(256) void k_probe_setup(ProbeDev d, uint32_t count, uint32_t T, uint32_t clampSteps,
                                                     uint32_t useBias) {
    const uint32_t sl = blockIdx.x;
    if (sl >= d.nslots) return;
    for (uint32_t i = threadIdx.x; i < d.ncol; i += 256u) d.spk[(size_t)sl * d.ncol + i] = 0u;
    for (uint32_t i = threadIdx.x; i < d.gw; i += 256u) d.gact[(size_t)sl * d.gw + i] = 0u;
    if (threadIdx.x == 0u) {
        d.st_t[sl] = 0u;
        d.st_T[sl] = sl < count ? T : 0u;
        d.st_clamp[sl] = clampSteps;
        d.st_bias[sl] = useBias;
        d.selDone[sl] = 0u;
        d.theta[2u * sl] = 0u;
        d.theta[2u * sl + 1u] = 0u;
    }
}

__global__ __launch_bounds__(256) void k_copy_words(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src,
                                                    size_t n) {
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256ull) dst[i] = src[i];
}
}  // namespace

hipError_t noalloc_scratch(int dev, uint32_t** out) {
    constexpr int MAXDEV = 64;
    static std::mutex mu;
    static uint32_t* addr[MAXDEV] = {};
    if (dev < 0 || dev >= MAXDEV) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        // the caller has made `dev` the current device: this is its copy of the array
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[dev] = static_cast<uint32_t*>(p);
    }
    *out = addr[dev];
    return hipSuccess;
}

hipError_t launch_probe_setup(const ProbeDev& d, uint32_t count, uint32_t T, uint32_t clampSteps, uint32_t useBias,
                              hipStream_t s) {
    k_probe_setup<<<d.nslots, 256, 0, s>>>(d, count, T, clampSteps, useBias);
    return hipGetLastError();
}

hipError_t launch_copy_words(uint32_t* dst, const uint32_t* src, size_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    k_copy_words<<<(uint32_t)std::min<size_t>((n + 255) / 256, 4096), 256, 0, s>>>(dst, src, n);
    return hipGetLastError();
}
}  // namespace axon
