// hx_ntt_kernels.hpp — NTT device code shared by hx_ntt.hip and hx_keyswitch.hip:
// plan layout, pass arguments, the generic and register-radix-16 pass kernels,
// and the host helpers that configure and launch a pass.  Everything below
// lives in an anonymous namespace, so each translation unit that includes
// this header gets its own copy of the kernels.
#pragma once
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <string>

#include "hx_internal.hpp"

struct LimbInfo {
    uint64_t q;
    uint64_t d;      // Möller–Granlund normalised divisor
    uint64_t v;      // and its reciprocal
    uint32_t s;
    uint32_t small;  // 1: q < 2^31, butterflies in 32-bit arithmetic
    uint64_t minv;   // m^-1 mod q (inverse plans)
    uint64_t minv_p; // its Shoup companion (32- or 64-bit)
    uint64_t wm;     // inv_twiddles[t, 1] * m^-1 mod q: the last inverse stage with the scaling folded in
    uint64_t wm_p;   // its Shoup companion
    uint64_t qni;    // -q^-1 mod 2^64 for odd q (Montgomery reduction), 0 for even q
};

struct hx_ntt_plan {
    int device;
    int inverse;
    int dtype;
    int logm;
    int64_t m;
    int64_t k;
    int perm_kind;        // PERM_* below
    int perm_bitrev;      // perm is the bit reversal of log m bits (a PERM_MAP that is its own inverse)
    // Every column runs the reference-semantics kernel (k_ntt_ref): the tables
    // hold entries outside [0, q), or a modulus is so large that the
    // reference's T-width sums wrap even for reduced inputs.
    int ref_all;
    void* mem;            // single device allocation (stream-ordered pool)
    LimbInfo* limbs;      // [k]
    uint64_t* tw;         // [k][m][2]: (w mod q, companion)
    int32_t* rowmap;      // [m] (see perm_kind)
    int64_t* raw_tw;      // [k][m] the table as given (ref_all plans only)
    int64_t* raw_minv;    // [k] m_inv as given (ref_all inverse plans only)
    // Intermediates between passes: one workspace per stream for eager work,
    // one per captured execution (see hx::Workspaces).
    hx::Workspaces* wss;
};

namespace hx {
namespace {

enum : int {
    PERM_IDENTITY = 0,
    PERM_MAP = 1,          // a permutation: rowmap[perm[u]] = u (fused)
    PERM_GATHER_FWD = 2,   // forward, not a permutation: out[u] = x[rowmap[u]] (rowmap = perm)
    PERM_SCATTER_INV = 3,  // inverse, not a permutation: x[j] = a[rowmap[j]] or old result if -1
};

constexpr int kThreads = 256;

__device__ __forceinline__ ModInfo limb_modinfo(const LimbInfo& L) {
    ModInfo mi;
    mi.M = L.q;
    mi.d = L.d;
    mi.v = L.v;
    mi.s = L.s;
    mi.pad = 0;
    return mi;
}

template <typename T>
__device__ __forceinline__ int64_t ldv(const void* base, int64_t off) {
    return (int64_t)((const T*)base)[off];
}

__device__ __forceinline__ int64_t ld_elem(const void* base, int64_t off, int elem32) {
    return elem32 ? (int64_t)((const int32_t*)base)[off] : ((const int64_t*)base)[off];
}

// Exact residue of a signed value (inputs of the NTT are non-negative; a
// negative one is reduced mathematically).
__device__ __forceinline__ uint64_t residue(const ModInfo& mi, int64_t x) {
    if (x >= 0) return reduce64(mi, (uint64_t)x);
    uint64_t r = reduce64(mi, uabs64(x));
    return r ? mi.M - r : 0;
}

// ------------------------------------------------------------ plan kernels
struct PlanSrc {
    const void* ptr;
    int64_t s0, s1;
    int32_t elem32;
};

// flags[5]: table entries outside [0, q); flags[6]: moduli above qmax (the
// reference's T-width sums of reduced values would wrap)
__global__ void k_plan_limbs(PlanSrc p, PlanSrc minv, int has_minv, int64_t k, int64_t qmax, LimbInfo* limbs,
                             unsigned long long* flags) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= k) return;
    int64_t q = ld_elem(p.ptr, t * p.s0, p.elem32);
    LimbInfo L;
    memset(&L, 0, sizeof L);
    if (q <= 1) {
        atomicAdd(&flags[0], 1ull);
        q = 2;  // keeps the table well-formed; the plan is rejected on the host
    }
    if (q > qmax) atomicAdd(&flags[6], 1ull);
    ModInfo mi = make_modinfo((uint64_t)q);
    L.q = (uint64_t)q;
    L.d = mi.d;
    L.v = mi.v;
    L.s = mi.s;
    L.small = q < (int64_t(1) << 31);
    L.qni = (q & 1) ? neg_inv64((uint64_t)q) : 0;
    if (has_minv) {
        const int64_t raw = ld_elem(minv.ptr, t * minv.s0, minv.elem32);
        if (raw < 0 || raw >= q) atomicAdd(&flags[5], 1ull);
        uint64_t mv = residue(mi, raw);
        L.minv = mv;
        L.minv_p = L.small ? divq128(mi, 0, mv << 32) : shoup_companion(mi, mv);
    }
    limbs[t] = L;
}

__global__ void k_plan_twiddles(PlanSrc tw, int64_t k, int64_t m, LimbInfo* limbs, int inverse, uint64_t* out,
                                unsigned long long* flags) {
    int64_t n = k * m;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t t = e / m, i = e - t * m;
        const LimbInfo& L = limbs[t];  // one thread per limb writes wm, wm_p; the others read q, d, v, s, small
        ModInfo mi = limb_modinfo(L);
        const int64_t raw = ld_elem(tw.ptr, t * tw.s0 + i * tw.s1, tw.elem32);
        if (raw < 0 || (uint64_t)raw >= L.q) atomicAdd(&flags[5], 1ull);
        uint64_t w = residue(mi, raw);
        uint64_t wp = L.small ? divq128(mi, 0, w << 32) : shoup_companion(mi, w);
        out[2 * e] = w;
        out[2 * e + 1] = wp;
        if (inverse && i == 1) {
            // the last Gentleman-Sande stage uses inv_twiddles[t, 1] only
            const uint64_t wm = mulmod(mi, w, L.minv);
            limbs[t].wm = wm;
            limbs[t].wm_p = L.small ? divq128(mi, 0, wm << 32) : shoup_companion(mi, wm);
        }
    }
}

__global__ void k_plan_perm_count(PlanSrc perm, int64_t m, int32_t* counts, unsigned long long* flags) {
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < m; u += (int64_t)gridDim.x * blockDim.x) {
        int64_t v = ld_elem(perm.ptr, u * perm.s0, perm.elem32);
        if (v < 0 || v >= m) {
            atomicAdd(&flags[1], 1ull);
            continue;
        }
        if (v != u) atomicAdd(&flags[2], 1ull);
        atomicAdd(&counts[v], 1);
    }
}

__global__ void k_plan_perm_check(const int32_t* counts, int64_t m, unsigned long long* flags) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x)
        if (counts[j] != 1) atomicAdd(&flags[3], 1ull);
}

// counts the entries of perm that differ from the bit reversal of log m bits
__global__ void k_plan_bitrev(PlanSrc perm, int64_t m, int logm, unsigned long long* flags) {
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < m; u += (int64_t)gridDim.x * blockDim.x) {
        const int64_t br = logm ? (int64_t)(__brev((uint32_t)u) >> (32 - logm)) : 0;
        if (ld_elem(perm.ptr, u * perm.s0, perm.elem32) != br) atomicAdd(&flags[4], 1ull);
    }
}

// mode 0: map[perm[u]] = u; mode 1: map[u] = perm[u]; mode 2: map[perm[u]] = max u (map pre-filled with -1)
__global__ void k_plan_map(PlanSrc perm, int64_t m, int32_t* map, int mode) {
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < m; u += (int64_t)gridDim.x * blockDim.x) {
        int32_t v = (int32_t)ld_elem(perm.ptr, u * perm.s0, perm.elem32);
        if (mode == 0) map[v] = (int32_t)u;
        else if (mode == 1) map[u] = v;
        else atomicMax(&map[v], (int32_t)u);
    }
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t val) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = val;
}

// ------------------------------------------------------------ pass kernel
// Column addressing: a column is (i, j, t); the three axes are listed in
// `order` (innermost first) with their sizes and strides for src and dst.
// The intermediate buffer uses rows of `ncols` uint64 with the column index
// innermost.
struct ColAddr {
    const void* ptr;
    int64_t row_stride;
    int64_t cstride[3];
    int32_t kind;  // 0: int64 tensor, 1: int32 tensor, 2: uint64 intermediate,
                   // 3: int64 tensor read as base-2^w digits (fused g_decomposition)
};

// Store epilogue of the key-switch's last pass for the serving loop's
// rotate-multiply-accumulate: instead of writing the key-switch output KS, the
// pass updates dst (the accumulator, [B, 2, m, k]) in place,
//     dst_c[row] = (dst_c[row] + rot_c * pt[row]) mod q_t,
//     rot_0 = (KS_0 + ct_0[j]) mod q_t,  rot_1 = KS_1,
// where row is the Galois row the key-switch writes coefficient j to, ct has
// dst's layout, and pt[row] = diags[b, diag, row mod M, t] ([B, M, M, k],
// contiguous per instance).  Each step is the reference op's own (modop_fast).
struct RmacArgs {
    int32_t on;
    const int64_t* ct;
    const int64_t* diags;
    int64_t diags_b, M, diag, L;  // L: the limbs of diags (its innermost dim)
};

struct PassArgs {
    ColAddr src;
    ColAddr dst;
    ColAddr alt;             // old result, read where a scatter leaves a row untouched
    int64_t ncols;
    int64_t cdim[3];
    int32_t limb_axis;       // which of the 3 column axes is t
    int32_t inverse;
    int32_t logm;
    int32_t b0;              // lowest index bit of the window
    int32_t H;               // window bits
    int32_t stage0;          // first global stage of this pass (h for forward, g for inverse)
    int32_t S_cols, S_bot;   // tile = S_cols columns x S_bot "bot" values
    int64_t bot_count;       // 2^b0
    int64_t n_bot_tiles, n_col_tiles;
    // 32-bit decoding of tile and column indices (set by pass_blocks, which
    // bounds the tile count and ncols below 2^31): divisors n_col_tiles,
    // n_bot_tiles, cdim[0], cdim[1]; S_cols = 2^log_scols
    FastDiv fd_ctiles, fd_btiles, fd_c0, fd_c1;
    int32_t log_scols;
    RmacArgs rm;             // store epilogue of the key-switch's last pass (rm.on)
    int32_t reduce_in;
    int32_t scale_out;
    int32_t lazy_out;        // radix-16 passes: leave 64-bit residues lazy ([0, 16q) forward, [0, 4q) inverse)
    int32_t lane_major;      // radix-16 passes: a sub-problem's lanes are adjacent threads (rows contiguous in memory)
    int32_t map_bitrev;      // src_map / dst_map is the bit reversal of logm bits: rows are computed, not loaded
    int32_t top_bitrev;      // radix-16 pass over the low bits (b0 = 0) of rows-contiguous columns whose row map is
                             // the bit reversal: tiles of 32 "top" values with adjacent bit reversals (r16_pass_brt)
    // digit source (src.kind == 3): column axis `gd_axis` is the digit index d,
    // the value loaded is digit d of x in base 2^gd_w; elements of digit 0 whose
    // value is not exhausted after gd_D digits are counted in gd_status
    int32_t gd_axis;
    int32_t gd_w;
    int32_t gd_D;
    unsigned long long* gd_status;
    const int32_t* dst_map_limb;  // per-limb row scatter on store: [k][m] (or null)
    const int32_t* src_map;  // row gather on load (or null)
    const int32_t* dst_map;  // row scatter on store (or null)
    int32_t src_map_alt;     // src_map entries of -1 read `alt` at the same row
    const LimbInfo* limbs;
    const uint64_t* tw;
    int64_t m;
    // Columns on which the reference's signed arithmetic gives other
    // representatives than the residues computed here (see k_ntt_ref):
    // colflag[col] is set by the first pass (flag_set) when a loaded value lies
    // above the limit of flag_kind, and the last pass (flag_skip) leaves the
    // result of a flagged column unwritten.  Null: no flags.
    uint32_t* colflag;
    int32_t flag_set;
    int32_t flag_skip;
    int32_t flag_kind;       // FLAG_* below
};

// Input limits above which a column is flagged (values compared as unsigned,
// so negative values are always above): the forward reference network is
// canonical for every non-negative input whose sums stay within int64; the
// inverse one for inputs in [0, q).
enum : int { FLAG_FWD64 = 0, FLAG_FWD32 = 1, FLAG_INV = 2 };

__device__ __forceinline__ uint64_t flag_limit(int kind, uint64_t q) {
    return kind == FLAG_INV ? q - 1 : (kind == FLAG_FWD32 ? ~0ull >> 1 : (~0ull >> 1) - q);
}

__device__ __forceinline__ int64_t col_offset(const ColAddr& c, const int64_t* coord) {
    return coord[0] * c.cstride[0] + coord[1] * c.cstride[1] + coord[2] * c.cstride[2];
}

__device__ __forceinline__ int64_t load_any(const ColAddr& c, int64_t off) {
    if (c.kind == 1) return ((const int32_t*)c.ptr)[off];
    return ((const int64_t*)c.ptr)[off];
}

__device__ __forceinline__ void store_any(const ColAddr& c, int64_t off, uint64_t v) {
    if (c.kind == 1) ((int32_t*)c.ptr)[off] = (int32_t)v;
    else ((uint64_t*)c.ptr)[off] = v;
}

// Workgroups are dispatched to the 8 XCDs round-robin by block index; each XCD
// gets a contiguous range of tiles so that tiles sharing cache lines meet in
// the same L2.
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t nb) {
    const int64_t nx = 8;
    int64_t x = b % nx, i = b / nx;
    int64_t per = nb / nx, rem = nb % nx;
    return x * per + (x < rem ? x : rem) + i;
}

// Per-tile bookkeeping shared by the pass kernels (one entry per sub-problem;
// N = 64 for the generic pass, 32 for the radix-16 pass).
template <int N>
struct TileSharedN {
    int64_t src_off[N];
    int64_t dst_off[N];
    int64_t alt_off[N];
    int64_t jbase[N];
    int64_t col[N];
    int32_t limb[N];
    int32_t digit[N];
    int32_t valid[N];
    int32_t bad[N];   // flag_set: a value of this sub-problem is above the flag limit
    int32_t skip[N];  // flag_skip: the column was flagged by an earlier pass
};
using TileShared = TileSharedN<64>;

// The first pass found a value of column `col` above the flag limit.  In a
// single pass the sub-problem is the whole column, so its own result is
// withheld too (bad); stores of the same value from several threads are benign.
template <int N>
__device__ __forceinline__ void mark_flagged(const PassArgs& P, TileSharedN<N>& T, int s) {
    T.bad[s] = 1;
    P.colflag[T.col[s]] = 1u;
}

// Fills T for tile `blk` of `nblk` (a tile index at or beyond nblk gives a
// tile of no valid sub-problem); tid is the thread's index within the threads
// that serve the tile.  Returns the index bits above the window, the same for
// every sub-problem of the tile.
template <int N>
__device__ __forceinline__ int64_t tile_setup(const PassArgs& P, int S, TileSharedN<N>& T, int64_t blk, int64_t nblk,
                                              int tid) {
    const bool in_range = blk < nblk;
    int64_t tile = in_range ? xcd_tile(blk, nblk) : 0;
    const int64_t col_tile = tile % P.n_col_tiles;
    tile /= P.n_col_tiles;
    const int64_t bot_tile = tile % P.n_bot_tiles;
    const int64_t top = tile / P.n_bot_tiles;
    if (tid < S) {
        int s = tid;
        int cs = s % P.S_cols, bs = s / P.S_cols;
        int64_t col = col_tile * P.S_cols + cs;
        int64_t bot = bot_tile * P.S_bot + bs;
        int valid = in_range && col < P.ncols && bot < P.bot_count;
        int64_t c = valid ? col : 0;
        int64_t coord[3];
        coord[0] = c % P.cdim[0];
        c /= P.cdim[0];
        coord[1] = c % P.cdim[1];
        coord[2] = c / P.cdim[1];
        // the intermediate buffer addresses columns by their linear index
        int64_t lcoord[3] = {valid ? col : 0, 0, 0};
        T.src_off[s] = P.src.kind == 2 ? col_offset(P.src, lcoord) : col_offset(P.src, coord);
        T.dst_off[s] = P.dst.kind == 2 ? col_offset(P.dst, lcoord) : col_offset(P.dst, coord);
        T.alt_off[s] = P.alt.ptr ? col_offset(P.alt, coord) : 0;
        T.limb[s] = (int32_t)coord[P.limb_axis];
        T.digit[s] = P.src.kind == 3 ? (int32_t)coord[P.gd_axis] : 0;
        T.valid[s] = valid;
        T.jbase[s] = (top << (P.b0 + P.H)) + (valid ? bot : 0);
        T.col[s] = valid ? col : 0;
        T.bad[s] = 0;
        T.skip[s] = (P.colflag && P.flag_skip && valid) ? (int32_t)P.colflag[col] : 0;
    }
    __syncthreads();
    return top;
}

// Coalesced load of the tile into LDS: consecutive threads take consecutive
// sub-problems of one window row.  `slot(s, w)` gives the LDS index.
template <typename Slot>
__device__ __forceinline__ void tile_load(const PassArgs& P, int S, int W, TileShared& T, uint64_t* xs, Slot slot) {
    const bool flags = P.colflag && P.flag_set;
    for (int e = threadIdx.x; e < S * W; e += blockDim.x) {
        int s = e % S, w = e / S;
        if (!T.valid[s]) continue;
        int64_t j = T.jbase[s] + ((int64_t)w << P.b0);
        int64_t x;
        if (P.src_map) {
            int32_t r = P.src_map[j];
            if (r >= 0) x = load_any(P.src, T.src_off[s] + (int64_t)r * P.src.row_stride);
            else x = load_any(P.alt, T.alt_off[s] + j * P.alt.row_stride);
        } else {
            x = load_any(P.src, T.src_off[s] + j * P.src.row_stride);
        }
        uint64_t v = (uint64_t)x;
        if (P.reduce_in) {
            LimbInfo L = P.limbs[T.limb[s]];
            if (x < 0 || (uint64_t)x >= L.q) v = residue(limb_modinfo(L), x);
        }
        if (flags && (uint64_t)x > flag_limit(P.flag_kind, P.limbs[T.limb[s]].q)) mark_flagged(P, T, s);
        xs[slot(s, w)] = v;
    }
    __syncthreads();
}

template <typename Slot>
__device__ __forceinline__ void tile_store(const PassArgs& P, int S, int W, const TileShared& T, const uint64_t* xs,
                                           Slot slot) {
    for (int e = threadIdx.x; e < S * W; e += blockDim.x) {
        int s = e % S, w = e / S;
        if (!T.valid[s]) continue;
        if (P.flag_skip && (T.skip[s] || T.bad[s])) continue;
        int64_t j = T.jbase[s] + ((int64_t)w << P.b0);
        int64_t row = P.dst_map ? (int64_t)P.dst_map[j] : j;
        uint64_t v = xs[slot(s, w)];
        if (P.scale_out) {
            const LimbInfo& L = P.limbs[T.limb[s]];
            if (L.small) v = shoup_mul32((uint32_t)v, (uint32_t)L.minv, (uint32_t)L.minv_p, (uint32_t)L.q);
            else v = shoup_mul(v, L.minv, L.minv_p, L.q);
        }
        store_any(P.dst, T.dst_off[s] + row * P.dst.row_stride, v);
    }
}

// Generic pass: any window size; one LDS round trip and barrier per stage.
__global__ __launch_bounds__(kThreads) void k_ntt_pass(PassArgs P) {
    extern __shared__ uint64_t xs[];  // [S][W + 1]
    __shared__ TileShared T;
    const int W = 1 << P.H;
    const int WP = W + 1;
    const int S = P.S_cols * P.S_bot;
    auto slot = [WP](int s, int w) { return s * WP + w; };
    tile_setup(P, S, T, blockIdx.x, gridDim.x, threadIdx.x);
    tile_load(P, S, W, T, xs, slot);

    const int half_count = W >> 1;
    for (int st = 0; st < P.H; ++st) {
        // forward: window bit H-1-st; inverse: window bit st
        const int bit = P.inverse ? st : (P.H - 1 - st);
        const int half = 1 << bit;
        const int g = P.stage0 + st;  // global stage
        for (int bf = threadIdx.x; bf < S * half_count; bf += blockDim.x) {
            int s = bf / half_count, i = bf - s * half_count;
            if (!T.valid[s]) continue;
            int lo = ((i >> bit) << (bit + 1)) | (i & (half - 1));
            int hi = lo + half;
            int64_t jlo = T.jbase[s] + ((int64_t)lo << P.b0);
            int64_t tw_idx = P.inverse ? ((int64_t(1) << (P.logm - 1 - g)) + (jlo >> (g + 1)))
                                       : ((int64_t(1) << g) + (jlo >> (P.logm - g)));
            const int limb = T.limb[s];
            const uint64_t* twp = P.tw + 2 * ((int64_t)limb * P.m + tw_idx);
            uint64_t w = twp[0], wp = twp[1];
            const LimbInfo& L = P.limbs[limb];
            uint64_t q = L.q;
            uint64_t u = xs[s * WP + lo], v = xs[s * WP + hi];
            uint64_t nu, nv;
            if (L.small) {
                uint32_t q32 = (uint32_t)q;
                if (!P.inverse) {
                    uint32_t t = shoup_mul32((uint32_t)v, (uint32_t)w, (uint32_t)wp, q32);
                    uint32_t a = (uint32_t)u + t;
                    nu = a >= q32 ? a - q32 : a;
                    nv = (uint32_t)u >= t ? (uint32_t)u - t : (uint32_t)u + q32 - t;
                } else {
                    uint32_t a = (uint32_t)u + (uint32_t)v;
                    nu = a >= q32 ? a - q32 : a;
                    uint32_t d = (uint32_t)u >= (uint32_t)v ? (uint32_t)u - (uint32_t)v : (uint32_t)u + q32 - (uint32_t)v;
                    nv = shoup_mul32(d, (uint32_t)w, (uint32_t)wp, q32);
                }
            } else {
                if (!P.inverse) {
                    uint64_t t = shoup_mul(v, w, wp, q);
                    nu = addmod(u, t, q);
                    nv = submod(u, t, q);
                } else {
                    nu = addmod(u, v, q);
                    nv = shoup_mul(submod(u, v, q), w, wp, q);
                }
            }
            xs[s * WP + lo] = nu;
            xs[s * WP + hi] = nv;
        }
        __syncthreads();
    }
    tile_store(P, S, W, T, xs, slot);
}

// ---------------------------------------------------------------------------
// Butterfly arithmetic of the register-radix-16 passes.
//
// Arithmetic per limb (uniform over a sub-problem):
//   MODE_S32  q < 2^31: 32-bit Shoup, residues kept in [0, q);
//   MODE_L64  q < 2^60: Harvey's lazy butterflies (D. Harvey, "Faster
//             arithmetic for number-theoretic transforms", J. Symb. Comp. 60,
//             2014) on the three-product Shoup remainder t in [0, 4q)
//             (shoup_lazy3_nq).  Forward: the upper input u is reduced (one
//             conditional subtraction of 8q) only at the even stages of a
//             window (RED), which keeps every value below 16q <= 2^64:
//                 RED:  u < 16q -> u < 8q,  outputs u + t, u - t + 4q < 12q;
//                 else: u < 12q,            outputs < 16q.
//             A pass starts at an even stage, so any lazy input below 16q is
//             valid.  Inverse values stay in [0, 4q).  Both are reduced when
//             leaving the transform;
//   MODE_F64  q < 2^63: 64-bit Shoup, residues kept in [0, q).
enum : int { MODE_S32 = 0, MODE_L64 = 1, MODE_F64 = 2 };

__device__ __forceinline__ int limb_mode(const LimbInfo& L) {
    return L.small ? MODE_S32 : (L.q < (uint64_t(1) << 60) ? MODE_L64 : MODE_F64);
}

// RED: reduce the forward butterfly's upper input (MODE_L64; see above)
template <int MODE, bool INV, bool RED = true>
__device__ __forceinline__ void butterfly(uint64_t& x0, uint64_t& x1, uint64_t w, uint64_t wp, uint64_t q) {
    if (MODE == MODE_S32) {
        uint32_t q32 = (uint32_t)q, u = (uint32_t)x0, v = (uint32_t)x1;
        if (!INV) {
            uint32_t t = shoup_mul32(v, (uint32_t)w, (uint32_t)wp, q32);
            uint32_t a = u + t;
            x0 = a >= q32 ? a - q32 : a;
            x1 = u >= t ? u - t : u + q32 - t;
        } else {
            uint32_t a = u + v;
            x0 = a >= q32 ? a - q32 : a;
            uint32_t d = u >= v ? u - v : u + q32 - v;
            x1 = shoup_mul32(d, (uint32_t)w, (uint32_t)wp, q32);
        }
    } else if (MODE == MODE_L64) {
        if (!INV) lazy_fwd_bfly<RED>(x0, x1, w, wp, q);
        else lazy_inv_bfly(x0, x1, w, wp, q);
    } else {
        if (!INV) {
            uint64_t t = shoup_mul(x1, w, wp, q);
            uint64_t u = x0;
            x0 = addmod(u, t, q);
            x1 = submod(u, t, q);
        } else {
            uint64_t u = x0, v = x1;
            x0 = addmod(u, v, q);
            x1 = shoup_mul(submod(u, v, q), w, wp, q);
        }
    }
}


// ---------------------------------------------------------------------------
// Register-radix-16 pass for windows of H = 6..8 bits (the NTT passes).
//
// A sub-problem (2^H points) is held by TL = 2^nb lanes (nb = H - 4) with 16
// values each, in one of two layouts:
//   layout A, window bits [nb, H) inside a lane:  w = e << nb | lane
//   layout B, window bits [0, nb) inside a lane:  w = (e >> nb) << 2nb | lane << nb | (e mod 2^nb)
// The forward pass loads straight from global memory into layout A, runs the
// four stages over the top window bits, moves the values through LDS into
// layout B, runs the remaining stages and stores from registers; the inverse
// pass does the same in the opposite order.  Thread t serves sub-problem
// t mod S and lane t / S, so a load instruction covers S adjacent columns (or
// rows) of the tile, and 32 consecutive threads address 32 different LDS
// rows of odd length (no bank conflicts in either layout).
template <int H>
__device__ __forceinline__ int r16_wA(int e, int lane) {
    return (e << (H - 4)) | lane;
}

template <int H>
__device__ __forceinline__ int r16_wB(int e, int lane) {
    constexpr int nb = H - 4;
    return ((e >> nb) << (2 * nb)) | (lane << nb) | (e & ((1 << nb) - 1));
}

// The stages of one layout on the 16 registers of a lane.  The twiddle index
// of a butterfly is ubase(stage) + (w >> (b + 1)) for window bit b; in layout
// A that is ubase + (e >> (b - nb + 1)), in layout B ubase + lane part + a
// constant, so each stage needs one address and immediate offsets.  The
// twiddles of stage k + 1 are requested before the butterflies of stage k run,
// which hides their load latency behind one stage of arithmetic.
template <int H, bool INV, int LAYOUT>
struct R16Stage {
    static constexpr int nb = H - 4;
    __device__ static constexpr int bit(int k) {  // forward: high bits first; inverse: low bits first
        return LAYOUT == 0 ? (INV ? nb + k : H - 1 - k) : (INV ? k : nb - 1 - k);
    }
    __device__ static constexpr int ebit(int k) { return LAYOUT == 0 ? bit(k) - nb : bit(k); }
    // twiddle offset of the butterfly whose lower register is e
    __device__ static constexpr int toff(int k, int e) {
        return LAYOUT == 0 ? (e >> (bit(k) - nb + 1))
                           : (((e >> nb) << (2 * nb - bit(k) - 1)) | ((e & ((1 << nb) - 1)) >> (bit(k) + 1)));
    }
    // the 8 twiddles of stage k, in butterfly order (equal addresses are loaded once)
    __device__ static __forceinline__ void load(ulonglong2 (&tw)[8], int k, int lane, const ulonglong2* twrow,
                                                int64_t top, int logm, int g0) {
        const int b = bit(k);
        const int st = INV ? b : H - 1 - b;  // stage within the window
        const int g = g0 + st;               // global stage
        const int64_t ubase = INV ? ((int64_t(1) << (logm - 1 - g)) + (top << (H - 1 - b)))
                                  : ((int64_t(1) << g) + (top << st));
        const ulonglong2* tp = twrow + ubase + (LAYOUT == 1 ? ((int64_t)lane << (nb - b - 1)) : 0);
        int i = 0;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            if (e & (1 << ebit(k))) continue;
            tw[i++] = tp[toff(k, e)];
        }
    }
};

// Last inverse stage with the m^-1 scaling folded in: (u + v) * m^-1 and
// (u - v) * (inv_twiddles[t, 1] * m^-1), fully reduced.
template <int MODE>
__device__ __forceinline__ void butterfly_last_scaled(uint64_t& x0, uint64_t& x1, const LimbInfo& L) {
    const uint64_t q = L.q;
    if (MODE == MODE_S32) {
        const uint32_t q32 = (uint32_t)q, u = (uint32_t)x0, v = (uint32_t)x1;
        x0 = shoup_mul32(u + v, (uint32_t)L.minv, (uint32_t)L.minv_p, q32);
        x1 = shoup_mul32(u - v + q32, (uint32_t)L.wm, (uint32_t)L.wm_p, q32);
    } else {
        const uint64_t lift = MODE == MODE_L64 ? q << 2 : q;  // inputs below 4q (lazy) or below q
        const uint64_t u = x0, v = x1;
        x0 = shoup_mul(u + v, L.minv, L.minv_p, q);
        x1 = shoup_mul(u - v + lift, L.wm, L.wm_p, q);
    }
}

// fold: the last stage is the transform's last inverse stage, scaled by m^-1
// (a uniform flag; only the last stage has two variants)
template <int H, bool INV, int MODE, int LAYOUT>
__device__ __forceinline__ void r16_stages(uint64_t (&x)[16], int lane, const ulonglong2* twrow, int64_t top,
                                           int logm, int g0, const LimbInfo& L, bool fold) {
    using St = R16Stage<H, INV, LAYOUT>;
    constexpr int nst = LAYOUT == 0 ? 4 : H - 4;
    const uint64_t q = L.q;
    ulonglong2 cur[8], nxt[8];
    __builtin_amdgcn_sched_barrier(0);  // the first twiddles are not hoisted into the caller's loads
    St::load(cur, 0, lane, twrow, top, logm, g0);
#pragma unroll
    for (int k = 0; k < nst; ++k) {
        // the inverse's first stage already holds 8 twiddles: its successor is
        // loaded after it (keeps the kernel at 4 waves per SIMD)
        const bool prefetch = k + 1 < nst && !(INV && k == 0);
        if (prefetch) {
            St::load(nxt, k + 1, lane, twrow, top, logm, g0);
            __builtin_amdgcn_sched_barrier(0);  // keep the requests ahead of this stage's arithmetic
        }
        const int eb = St::ebit(k);
        if (INV && LAYOUT == 0 && k == nst - 1 && fold) {
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                if (e & (1 << eb)) continue;
                butterfly_last_scaled<MODE>(x[e], x[e | (1 << eb)], L);
            }
        } else {
            // the window's stage is k (layout A) or 4 + k (layout B): even for even k
            int i = 0;
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                if (e & (1 << eb)) continue;
                if (k & 1) butterfly<MODE, INV, false>(x[e], x[e | (1 << eb)], cur[i].x, cur[i].y, q);
                else butterfly<MODE, INV, true>(x[e], x[e | (1 << eb)], cur[i].x, cur[i].y, q);
                ++i;
            }
        }
        if (prefetch) {
#pragma unroll
            for (int j = 0; j < 8; ++j) cur[j] = nxt[j];
        } else if (k + 1 < nst) {
            St::load(cur, k + 1, lane, twrow, top, logm, g0);
        }
    }
}

template <int H, bool INV, int LAYOUT>
__device__ __forceinline__ void r16_stages_any(uint64_t (&x)[16], int lane, const ulonglong2* twrow, int64_t top,
                                               int logm, int g0, const LimbInfo& L, int mode, bool fold = false) {
    if (mode == MODE_L64) r16_stages<H, INV, MODE_L64, LAYOUT>(x, lane, twrow, top, logm, g0, L, fold);
    else if (mode == MODE_S32) r16_stages<H, INV, MODE_S32, LAYOUT>(x, lane, twrow, top, logm, g0, L, fold);
    else r16_stages<H, INV, MODE_F64, LAYOUT>(x, lane, twrow, top, logm, g0, L, fold);
}

template <int H>
__device__ __forceinline__ int r16_w(int layout, int e, int lane) {
    return layout == 0 ? r16_wA<H>(e, lane) : r16_wB<H>(e, lane);
}

// w = r16_wl(layout, lane) + r16_wc(layout, e): the per-lane part of a window
// index and the part fixed by the register number, so that an address is one
// per-lane base plus a uniform offset per register.
template <int H>
__device__ __forceinline__ int r16_wl(int layout, int lane) {
    return layout == 0 ? lane : lane << (H - 4);
}

template <int H>
__device__ __forceinline__ int r16_wc(int layout, int e) {
    constexpr int nb = H - 4;
    return layout == 0 ? e << nb : ((e >> nb) << (2 * nb)) + (e & ((1 << nb) - 1));
}

// LDS slot of window index w.  LM (lane-major thread order, for layouts whose
// rows are contiguous): one pad slot every 16 and rows of W + W/16, so that 16
// adjacent lanes of either layout and the sub-problems of a half-wave hit
// different banks; otherwise odd rows of W + 1 (32 adjacent threads are 32
// different sub-problems).
template <int H, bool LM>
struct R16Lds {
    static constexpr int W = 1 << H;
    static constexpr int SP = LM ? W + W / 16 : W + 1;
    __device__ static constexpr int slot(int w) { return LM ? w + (w >> 4) : w; }
};

// Sub-problem s of tile `blk` of a radix-16 pass, decoded by each of its
// threads on its own (no barrier before the loads): column, window base row,
// offsets, limb and digit, and whether an earlier pass flagged the column.
struct R16Sub {
    int64_t src_off, dst_off, alt_off, jbase, col, top;
    int32_t limb, digit, valid, skip;
    int32_t cc0, cc2;  // column coordinates 0 and 2
};

// Sub-problem s of tile `tile` (already mapped to the tile order; in_range
// false gives a sub-problem that is not valid).
__device__ __forceinline__ R16Sub r16_sub_at(const PassArgs& P, uint32_t tile, bool in_range, int s) {
    R16Sub u;
    const uint32_t tq = fdiv(tile, P.fd_ctiles);
    const uint32_t col_tile = tile - tq * P.fd_ctiles.d;
    const uint32_t top = fdiv(tq, P.fd_btiles);
    const uint32_t bot_tile = tq - top * P.fd_btiles.d;
    const uint32_t cs = (uint32_t)s & ((1u << P.log_scols) - 1u), bs = (uint32_t)s >> P.log_scols;
    const int64_t col = ((int64_t)col_tile << P.log_scols) + cs;
    const int64_t bot = (int64_t)bot_tile * P.S_bot + bs;
    const bool valid = in_range && col < P.ncols && bot < P.bot_count;
    const uint32_t c = valid ? (uint32_t)col : 0u;
    const uint32_t c1 = fdiv(c, P.fd_c0);
    const uint32_t c2 = fdiv(c1, P.fd_c1);
    const int64_t coord[3] = {(int64_t)(c - c1 * P.fd_c0.d), (int64_t)(c1 - c2 * P.fd_c1.d), (int64_t)c2};
    // the intermediate buffer addresses columns by their linear index
    const int64_t lcoord[3] = {(int64_t)c, 0, 0};
    u.src_off = P.src.kind == 2 ? col_offset(P.src, lcoord) : col_offset(P.src, coord);
    u.dst_off = P.dst.kind == 2 ? col_offset(P.dst, lcoord) : col_offset(P.dst, coord);
    u.alt_off = P.alt.ptr ? col_offset(P.alt, coord) : 0;
    u.limb = (int32_t)coord[P.limb_axis];
    u.digit = P.src.kind == 3 ? (int32_t)coord[P.gd_axis] : 0;
    u.cc0 = (int32_t)coord[0];
    u.cc2 = (int32_t)coord[2];
    u.valid = valid;
    u.top = top;
    u.jbase = ((int64_t)top << (P.b0 + P.H)) + (valid ? bot : 0);
    u.col = c;
    u.skip = (P.colflag && P.flag_skip && valid) ? (int32_t)P.colflag[c] : 0;
    return u;
}

__device__ __forceinline__ R16Sub r16_sub(const PassArgs& P, int64_t blk, int64_t nblk, int s) {
    const bool in_range = blk < nblk;
    return r16_sub_at(P, in_range ? (uint32_t)xcd_tile(blk, nblk) : 0u, in_range, s);
}

// Per-workgroup bookkeeping of a radix-16 tile: the combined flags of a
// sub-problem's threads (a pass that both flags and withholds columns), and
// the destination rows of the tile's window indices when the pass scatters
// through a row map shared by all columns (loaded with the data, read at the
// store).
struct R16Shared {
    int32_t bad[32];
    int32_t rows[256];
};

// Tile `blk` of a radix-16 pass's `nblk` tiles, served by the 32 * 2^(H-4)
// threads tid = 0.. of a workgroup with tile bookkeeping T and an LDS block xs
// of [S][SP] slots (S = 32 or 16 sub-problems).  Every thread of the workgroup
// runs the same barriers.
// values per load group of the accumulate epilogue: four keeps the H = 8 instance
// within 128 VGPRs (four waves per SIMD) without spills
constexpr int kRmGroup = 4;

template <int H, bool INV, bool LM = false, bool RM = false, int S = 32>
__device__ __forceinline__ void r16_tile(const PassArgs& P, int64_t blk, int64_t nblk, int tid, R16Shared& T,
                                         uint64_t* xs) {
    static_assert(S == 16 || S == 32, "sub-problems per tile");
    constexpr int W = 1 << H;
    constexpr int SP = R16Lds<H, LM>::SP;
    constexpr int L_IN = INV ? 1 : 0;   // layout of the first stages
    constexpr int L_OUT = 1 - L_IN;
    constexpr int TL = W / 16;
    const int s = LM ? tid / TL : tid % S;
    const int lane = LM ? tid % TL : tid / S;
    const R16Sub u = r16_sub(P, blk, nblk, s);
    const int64_t top = u.top;
    // a pass that flags columns and also withholds their results (a single-pass
    // transform) combines the flags of a sub-problem's threads in T.bad,
    // cleared here and set after the exchange barrier
    const bool flag_here = P.colflag && P.flag_set && P.flag_skip;
    if (flag_here && lane == 0) T.bad[s] = 0;
    bool mybad = false;

    const bool valid = u.valid;
    const int limb = u.limb;
    const LimbInfo& L = P.limbs[limb];
    const uint64_t q = L.q;
    const int mode = limb_mode(L);
    const int64_t jbase = u.jbase;
    uint64_t x[16];

    // ---- load straight into the first layout
    if (valid) {
        const int es = P.src.kind == 1 ? 4 : 8;
        const int64_t wstride = (int64_t)es * (P.src.row_stride << P.b0);  // bytes per window row
        const char* base = (const char*)P.src.ptr + es * (u.src_off + jbase * P.src.row_stride) +
                           (int64_t)r16_wl<H>(L_IN, lane) * wstride;
        if (P.src.kind == 3) {
            // digit d of x: the reference's unsigned digit loop
            // (example_impl/g_decomposition_impl.cpp:52-56) in closed form
            const int d = u.digit;
            const int wb = P.gd_w;
            const uint64_t mask = wb ? (~0ull >> (64 - wb)) : 0;
            const int shift = wb * d;
            const int rest = wb * P.gd_D;
            unsigned long long over = 0;
            // all 16 loads are in flight before the first use: the reduction
            // branch below would otherwise keep the next load behind it
            int64_t xw[16];
#pragma unroll
            for (int e = 0; e < 16; ++e) xw[e] = *(const int64_t*)(base + r16_wc<H>(L_IN, e) * wstride);
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int64_t xv = xw[e];
                const uint64_t u = (uint64_t)xv;
                uint64_t v = (wb && shift < 64) ? ((u >> shift) & mask) : 0;
                if (d == 0) over += wb ? (rest < 64 && (u >> rest) != 0) : (xv > 0);
                if (P.reduce_in && v >= q) v = residue(limb_modinfo(L), (int64_t)v);
                x[e] = v;
            }
            if (over) atomicAdd(P.gd_status, over);
        } else {
            // raw values first (all 16 loads in flight), then reduction and flags
            int64_t xv[16];
            if (!P.src_map) {
                if (es == 8) {
#pragma unroll
                    for (int e = 0; e < 16; ++e) xv[e] = *(const int64_t*)(base + r16_wc<H>(L_IN, e) * wstride);
                } else {
#pragma unroll
                    for (int e = 0; e < 16; ++e) xv[e] = *(const int32_t*)(base + r16_wc<H>(L_IN, e) * wstride);
                }
            } else if (P.b0 == 0 && !P.src_map_alt && INV && es == 8) {
                // inverse row gather of the first pass: the 16 map entries of a lane
                // form runs of 2^(H-4) >= 4 consecutive rows, read as int4 (or
                // computed when the map is the bit reversal)
                int32_t r[16];
                if (P.map_bitrev) {
#pragma unroll
                    for (int e = 0; e < 16; ++e) {
                        const int64_t j = jbase + r16_wl<H>(L_IN, lane) + r16_wc<H>(L_IN, e);
                        r[e] = (int32_t)(__brev((uint32_t)j) >> (32 - P.logm));
                    }
                } else {
                    const int32_t* mp = P.src_map + jbase + r16_wl<H>(L_IN, lane);
#pragma unroll
                    for (int e = 0; e < 16; e += 4) {
                        const int4 v4 = *(const int4*)(mp + r16_wc<H>(L_IN, e));
                        r[e] = v4.x;
                        r[e + 1] = v4.y;
                        r[e + 2] = v4.z;
                        r[e + 3] = v4.w;
                    }
                }
                const char* colp = (const char*)P.src.ptr + 8 * u.src_off;
                const int64_t rs = 8 * P.src.row_stride;
#pragma unroll
                for (int e = 0; e < 16; ++e) xv[e] = *(const int64_t*)(colp + (int64_t)r[e] * rs);
            } else {
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const int64_t j = jbase + ((int64_t)r16_w<H>(L_IN, e, lane) << P.b0);
                    const int32_t r = P.src_map[j];
                    xv[e] = r >= 0 ? load_any(P.src, u.src_off + (int64_t)r * P.src.row_stride)
                                   : load_any(P.alt, u.alt_off + j * P.alt.row_stride);
                }
            }
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                uint64_t v = (uint64_t)xv[e];
                if (P.reduce_in && (xv[e] < 0 || v >= q)) v = residue(limb_modinfo(L), xv[e]);
                x[e] = v;
            }
            if (P.colflag && P.flag_set) {
                const uint64_t lim = flag_limit(P.flag_kind, q);
                bool bad = false;
#pragma unroll
                for (int e = 0; e < 16; ++e) bad |= (uint64_t)xv[e] > lim;
                if (bad) {
                    mybad = true;
                    P.colflag[u.col] = 1u;
                }
            }
        }
    } else {
#pragma unroll
        for (int e = 0; e < 16; ++e) x[e] = 0;
    }

    // destination rows of a row map shared by the tile's columns (S_bot = 1: one
    // window base row for the tile), requested behind the data; the inverse's
    // last pass (the key-switch's Galois scatter) is the pass that has one
    const bool rows_lds = INV && P.dst_map && !P.map_bitrev && !P.dst_map_limb && P.S_bot == 1;
    int32_t my_row = 0;
    if (rows_lds && tid < W) my_row = P.dst_map[r16_sub(P, blk, nblk, 0).jbase + ((int64_t)tid << P.b0)];

    const ulonglong2* twrow = (const ulonglong2*)(P.tw + 2 * (int64_t)limb * P.m);
    // ---- first stages, exchange, remaining stages
    if (valid) r16_stages_any<H, INV, L_IN>(x, lane, twrow, top, P.logm, P.stage0, L, mode);
    if (rows_lds && tid < W) T.rows[tid] = my_row;
    uint64_t* xs_s = xs + s * SP;
#pragma unroll
    for (int e = 0; e < 16; ++e) xs_s[R16Lds<H, LM>::slot(r16_w<H>(L_IN, e, lane))] = x[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; ++e) x[e] = xs_s[R16Lds<H, LM>::slot(r16_w<H>(L_OUT, e, lane))];
    bool withhold = u.skip != 0;
    if (flag_here) {  // uniform over the workgroup
        if (mybad) T.bad[s] = 1;
        __syncthreads();
        withhold = withhold || T.bad[s];
    }
    if (!valid) return;
    // the last pass of an inverse transform ends with the global stage logm - 1
    // (window bit H - 1, the last stage of layout A), where m^-1 is folded in
    r16_stages_any<H, INV, L_OUT>(x, lane, twrow, top, P.logm, P.stage0, L, mode, INV && P.scale_out);

    // ---- full reduction of lazy residues (scaled outputs are already reduced)
    if (!P.scale_out && !P.lazy_out && mode == MODE_L64) {
#pragma unroll
        for (int e = 0; e < 16; ++e) x[e] = lazy_reduce(x[e], q);
    }
    // a flagged column is written by the reference-semantics kernel instead
    if (P.flag_skip && withhold) return;

    // ---- store from registers
    const int es = P.dst.kind == 1 ? 4 : 8;
    const int64_t colbase = u.dst_off;
    if (!P.dst_map && !P.dst_map_limb) {
        const int64_t wstride = (int64_t)es * (P.dst.row_stride << P.b0);
        char* base = (char*)P.dst.ptr + es * (colbase + jbase * P.dst.row_stride) +
                     (int64_t)r16_wl<H>(L_OUT, lane) * wstride;
        if (es == 8) {
#pragma unroll
            for (int e = 0; e < 16; ++e)  // streaming: whole lines, read next by another launch
                __builtin_nontemporal_store(x[e], (uint64_t*)(base + r16_wc<H>(L_OUT, e) * wstride));
        } else {
#pragma unroll
            for (int e = 0; e < 16; ++e) *(int32_t*)(base + r16_wc<H>(L_OUT, e) * wstride) = (int32_t)x[e];
        }
        return;
    }
    const int32_t* map = P.dst_map_limb ? P.dst_map_limb + (int64_t)limb * P.m : P.dst_map;
    if constexpr (RM) {
        // rotate-multiply-accumulate into dst (the key-switch's last pass only)
        const ModInfo mi = limb_modinfo(L);
        const int64_t pt_base = (int64_t)u.cc2 * P.rm.diags_b + P.rm.diag * P.rm.M * P.rm.L + limb;
        const int64_t mmask = P.rm.M - 1;
        const int64_t* acc = (const int64_t*)P.dst.ptr;
        // the loads of G values are issued before their stores: a store may alias a later load as far as the
        // compiler knows, so loads interleaved with stores would wait one HBM round trip per value
        constexpr int G = kRmGroup;
#pragma unroll
        for (int h = 0; h < 16; h += G) {
            int64_t ao[G], av[G], cv[G], pv[G];
#pragma unroll
            for (int e = 0; e < G; ++e) {
                const int w = r16_w<H>(L_OUT, h + e, lane);
                const int64_t j = jbase + ((int64_t)w << P.b0);
                const int64_t row = P.map_bitrev ? (int64_t)(__brev((uint32_t)j) >> (32 - P.logm))
                                                 : (rows_lds ? (int64_t)T.rows[w] : (int64_t)map[j]);
                ao[e] = colbase + row * P.dst.row_stride;
                av[e] = acc[ao[e]];
                cv[e] = u.cc0 == 0 ? P.rm.ct[colbase + j * P.dst.row_stride] : 0;
                pv[e] = P.rm.diags[pt_base + (row & mmask) * P.rm.L];
            }
#pragma unroll
            for (int e = 0; e < G; ++e) {
                int64_t rot = (int64_t)x[h + e];
                if (u.cc0 == 0) rot = (int64_t)modop_fast(1, mi, rot, cv[e]);
                ((int64_t*)P.dst.ptr)[ao[e]] = (int64_t)modop_fast(1, mi, av[e], (int64_t)modop_fast(0, mi, rot, pv[e]));
            }
        }
        return;
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int64_t j = jbase + ((int64_t)r16_w<H>(L_OUT, e, lane) << P.b0);
        const int64_t row = P.map_bitrev ? (int64_t)(__brev((uint32_t)j) >> (32 - P.logm))
                                         : (rows_lds ? (int64_t)T.rows[r16_w<H>(L_OUT, e, lane)] : (int64_t)map[j]);
        char* ptr = (char*)P.dst.ptr + es * (colbase + row * P.dst.row_stride);
        if (es == 4) *(int32_t*)ptr = (int32_t)x[e];
        else *(uint64_t*)ptr = x[e];
    }
}

// One workgroup of a radix-16 pass: tile `blk` of the pass's `nblk` tiles.
template <int H, bool INV, bool LM = false, bool RM = false, int S = 32>
__device__ __forceinline__ void r16_pass(const PassArgs& P, int64_t blk, int64_t nblk) {
    extern __shared__ uint64_t xs[];    // [S][SP]
    __shared__ R16Shared T;
    r16_tile<H, INV, LM, RM, S>(P, blk, nblk, threadIdx.x, T, xs);
}

// RM: the inverse pass whose store is the rotate-multiply-accumulate epilogue
// (P.rm), an instance of its own so that the plain passes keep their register
// budget.  S: sub-problems per workgroup (S * 2^(H-4) threads).
template <int H, bool INV, bool LM, bool RM, int S>
__global__ __launch_bounds__(S << (H - 4), RM && H == 8 && S == 32 ? 4 : 1) void k_ntt_pass_r16(PassArgs P) {
    r16_pass<H, INV, LM, RM, S>(P, blockIdx.x, gridDim.x);
}


// The radix-16 pass over the low index bits (b0 = 0) of one rows-contiguous
// column whose row map is the bit reversal of log m bits: the forward pass
// stores x[j] at row brev(j), the inverse pass loads x[j] from row brev(j).  With
// j = top * 2^H + w, brev(j) = brev_H(w) * 2^(log m - H) + brev(top), so a tile
// of the 32 tops whose bit reversals are v0 .. v0 + 31 touches 32 adjacent
// permuted rows for every w.  The side of the pass at the permuted rows (the
// forward store, the inverse load) runs with the 32 sub-problems on adjacent
// threads; the other side with the 2^(H-4) lanes of one sub-problem on adjacent
// threads (adjacent rows); the LDS exchange between the two rounds switches
// the thread order.  The column, and so the limb, is the same for the tile;
// rows are adjacent elements on both sides (row stride 1).
template <int H, bool INV>
__device__ __forceinline__ void r16_pass_brt(const PassArgs& P, int64_t blk, int64_t nblk) {
    constexpr int W = 1 << H;
    constexpr int TL = W / 16;
    constexpr int S = 32;
    constexpr int SP = W + 1;
    constexpr int L_IN = INV ? 1 : 0;
    constexpr int L_OUT = 1 - L_IN;
    extern __shared__ uint64_t xs[];  // [S][SP]
    const int ntb = P.logm - H;       // bits of top
    const int64_t tile = xcd_tile(blk, nblk);
    const int64_t col = tile % P.ncols;
    const int64_t v0 = (tile / P.ncols) * S;  // bit reversals of the tile's tops: v0 .. v0 + 31
    int64_t c = col, coord[3];
    coord[0] = c % P.cdim[0];
    c /= P.cdim[0];
    coord[1] = c % P.cdim[1];
    coord[2] = c / P.cdim[1];
    const int64_t lcoord[3] = {col, 0, 0};
    const int64_t src_off = P.src.kind == 2 ? col_offset(P.src, lcoord) : col_offset(P.src, coord);
    const int64_t dst_off = P.dst.kind == 2 ? col_offset(P.dst, lcoord) : col_offset(P.dst, coord);
    const int limb = (int)coord[P.limb_axis];
    const LimbInfo& L = P.limbs[limb];
    const uint64_t q = L.q;
    const int mode = limb_mode(L);
    const ulonglong2* twrow = (const ulonglong2*)(P.tw + 2 * (int64_t)limb * P.m);
    const int tid = threadIdx.x;
    // thread orders: "rows" (lanes adjacent) and "perm" (sub-problems adjacent)
    const int s_r = tid / TL, lane_r = tid % TL;
    const int s_p = tid % S, lane_p = tid / S;
    const int s1 = INV ? s_p : s_r, lane1 = INV ? lane_p : lane_r;
    const int s2 = INV ? s_r : s_p, lane2 = INV ? lane_r : lane_p;
    const int64_t top1 = (int64_t)(__brev((uint32_t)(v0 + s1)) >> (32 - ntb));
    const int64_t top2 = (int64_t)(__brev((uint32_t)(v0 + s2)) >> (32 - ntb));
    const int es_in = P.src.kind == 1 ? 4 : 8, es_out = P.dst.kind == 1 ? 4 : 8;
    uint64_t x[16];
    // a column flagged by the first pass is written by the reference-semantics
    // kernel (uniform over the workgroup: the tile is one column)
    if (P.flag_skip && P.colflag[col]) return;

    // ---- load: forward rows top * W + w; inverse rows brev(top * W + w)
    // (rows are adjacent elements: one column pointer for the tile, 32-bit rows)
    const int64_t* col64 = (const int64_t*)P.src.ptr + src_off;
    const int32_t* col32 = (const int32_t*)P.src.ptr + src_off;
    uint32_t rows[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const uint32_t w = (uint32_t)r16_w<H>(L_IN, e, lane1);
        rows[e] = INV ? ((__brev(w) >> (32 - H)) << ntb) + (uint32_t)(v0 + s1) : ((uint32_t)top1 << H) + w;
    }
    // all 16 loads are in flight before the first reduction branch
    int64_t xv[16];
    if (es_in == 4) {
#pragma unroll
        for (int e = 0; e < 16; ++e) xv[e] = (int64_t)col32[rows[e]];
    } else {
#pragma unroll
        for (int e = 0; e < 16; ++e) xv[e] = col64[rows[e]];
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        uint64_t v = (uint64_t)xv[e];
        if (P.reduce_in && (xv[e] < 0 || v >= q)) v = residue(limb_modinfo(L), xv[e]);
        x[e] = v;
    }
    if (P.colflag && P.flag_set) {
        const uint64_t lim = flag_limit(P.flag_kind, q);
        bool bad = false;
#pragma unroll
        for (int e = 0; e < 16; ++e) bad |= (uint64_t)xv[e] > lim;
        if (bad) P.colflag[col] = 1u;
    }
    r16_stages_any<H, INV, L_IN>(x, lane1, twrow, top1, P.logm, P.stage0, L, mode);
#pragma unroll
    for (int e = 0; e < 16; ++e) xs[s1 * SP + r16_w<H>(L_IN, e, lane1)] = x[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; ++e) x[e] = xs[s2 * SP + r16_w<H>(L_OUT, e, lane2)];
    r16_stages_any<H, INV, L_OUT>(x, lane2, twrow, top2, P.logm, P.stage0, L, mode, INV && P.scale_out);
    if (!P.scale_out && !P.lazy_out && mode == MODE_L64) {
#pragma unroll
        for (int e = 0; e < 16; ++e) x[e] = lazy_reduce(x[e], q);
    }
    // ---- store: forward rows brev(top * W + w); inverse rows top * W + w
    uint64_t* out64 = (uint64_t*)P.dst.ptr + dst_off;
    int32_t* out32 = (int32_t*)P.dst.ptr + dst_off;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const uint32_t w = (uint32_t)r16_w<H>(L_OUT, e, lane2);
        const uint32_t row = INV ? ((uint32_t)top2 << H) + w : ((__brev(w) >> (32 - H)) << ntb) + (uint32_t)(v0 + s2);
        if (es_out == 4) out32[row] = (int32_t)x[e];
        else out64[row] = x[e];
    }
}

template <int H, bool INV>
__global__ __launch_bounds__(32 << (H - 4)) void k_ntt_pass_r16_brt(PassArgs P) {
    r16_pass_brt<H, INV>(P, blockIdx.x, gridDim.x);
}

// Index of the instance whose tile range [starts[i], starts[i + 1]) holds blk.
__device__ __forceinline__ int instance_of(const int64_t* starts, int count, int64_t blk) {
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= blk) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Tiles of a radix-16 pass (PassArgs prepared by configure_tile / pass_blocks).
__device__ __forceinline__ int64_t r16_tiles(const PassArgs& P) {
    return (int64_t(1) << (P.logm - P.b0 - P.H)) * P.n_bot_tiles * P.n_col_tiles;
}

// Passes of any window size H = 6..8 (one direction) in one launch of 512
// threads: workgroup b serves pass list[i] with b' = b - starts[i], and runs
// its 2^(8-H) tiles b' * 2^(8-H) .. as many groups of 32 * 2^(H-4) threads, so
// that every workgroup uses the same threads and about the same LDS
// (2^(8-H) * 32 * (2^H + 1) slots) whatever its pass.
template <bool INV>
__global__ __launch_bounds__(512, 4) void k_ntt_pass_r16_group(const PassArgs* list, const int64_t* starts,
                                                               int count) {
    extern __shared__ uint64_t xs[];
    __shared__ R16Shared T[4];
    const int i = instance_of(starts, count, blockIdx.x);
    const int64_t b = (int64_t)blockIdx.x - starts[i];
    const PassArgs& P = list[i];
    const int64_t ntiles = r16_tiles(P);
    const int H = P.H;
    const int sub = threadIdx.x >> (H + 1), tid = threadIdx.x & ((1 << (H + 1)) - 1);
    const int64_t tile = (b << (8 - H)) + sub;
    if (H == 8) r16_tile<8, INV>(P, tile, ntiles, tid, T[0], xs);
    else if (H == 7) r16_tile<7, INV>(P, tile, ntiles, tid, T[sub], xs + sub * 32 * ((1 << 7) + 1));
    else r16_tile<6, INV>(P, tile, ntiles, tid, T[sub], xs + sub * 32 * ((1 << 6) + 1));
}

// dst[u][col] = src[map[u]][col] (forward gather for non-permutation perms)
struct GatherArgs {
    ColAddr src;
    ColAddr dst;
    int64_t ncols;
    int64_t cdim[3];
    const int32_t* map;
    int64_t m;
    const uint32_t* colflag;  // flagged columns are left to the reference-semantics kernel
};

__global__ void k_row_gather(GatherArgs G) {
    int64_t n = G.m * G.ncols;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t u = e / G.ncols, col = e - u * G.ncols;
        if (G.colflag && G.colflag[col]) continue;
        int64_t c = col, coord[3];
        coord[0] = c % G.cdim[0];
        c /= G.cdim[0];
        coord[1] = c % G.cdim[1];
        coord[2] = c / G.cdim[1];
        int64_t lc[3] = {col, 0, 0};
        int64_t so = G.src.kind == 2 ? col_offset(G.src, lc) : col_offset(G.src, coord);
        int64_t dof = G.dst.kind == 2 ? col_offset(G.dst, lc) : col_offset(G.dst, coord);
        int64_t v = load_any(G.src, so + (int64_t)G.map[u] * G.src.row_stride);
        store_any(G.dst, dof + u * G.dst.row_stride, (uint64_t)v);
    }
}

// ---------------------------------------------------------------------------
// Reference semantics for flagged columns.
//
// The reference runs its butterflies on signed values with C++'s truncating
// `%` (example_impl/ntt_impl.cpp:101-108 forward, :151-157 inverse): sums in
// the element type T, products in the double width T_DP.  A negative input
// (and, for the inverse, one at or above q) therefore propagates negative
// representatives in (-q, q): congruent to what the passes compute, but not
// equal.  For every column flagged by the first pass this kernel replays the
// reference network step for step on the column's original input: one
// workgroup per column, the values in a global scratch row, one barrier per
// stage.  Sums wrap in 64 bits like the reference's int64 arithmetic; int32
// values are cut to 32 bits where the reference stores them in T.  The INTT's
// m^-1 scaling is exact (the reference multiplies in T, which overflows for
// q > ~2^31.5 in int64).  The kernel clears each flag it has served.
struct RefArgs {
    ColAddr src;            // a
    ColAddr old;            // the result's previous contents (rows an inverse scatter does not hit)
    ColAddr dst;            // result
    int64_t* scratch;       // [gridDim.x][m]
    uint32_t* colflag;      // [ncols]
    const int32_t* rowmap;  // the plan's row map (meaning per perm_kind)
    int32_t perm_kind;
    int32_t inverse;
    int32_t is32;
    int32_t logm;
    int64_t m, ncols;
    int64_t cdim[3];
    int32_t limb_axis;
    const LimbInfo* limbs;
    const uint64_t* tw;      // reduced table pairs, equal to the given table unless raw_tw is set
    const int64_t* raw_tw;   // [k][m] or null
    const int64_t* raw_minv; // [k] or null
};

__device__ __forceinline__ int64_t ref_t(int64_t x, int is32) { return is32 ? (int64_t)(int32_t)x : x; }

// static_cast<T>(T_DP(x) * T_DP(y) % T_DP(q)) for x, y of type T
__device__ __forceinline__ int64_t ref_mulmod(const ModInfo& mi, int is32, int64_t x, int64_t y) {
    if (is32) return cpp_rem(mi, x * y);  // |x|, |y| < 2^31: exact in int64; |result| < q < 2^31
    const bool neg = (x < 0) != (y < 0);
    const uint64_t ux = uabs64(x), uy = uabs64(y);
    const uint64_t r = reduce128(mi, reduce64(mi, mulhi64(ux, uy)), ux * uy);
    return apply_sign(neg && r != 0, r);
}

__global__ __launch_bounds__(256) void k_ntt_ref(RefArgs A) {
    const int64_t m = A.m, half = m >> 1;
    int64_t* x = A.scratch + (int64_t)blockIdx.x * m;
    for (int64_t col = blockIdx.x; col < A.ncols; col += gridDim.x) {
        if (!A.colflag[col]) continue;  // the same value for every thread of the workgroup
        int64_t c = col, coord[3];
        coord[0] = c % A.cdim[0];
        c /= A.cdim[0];
        coord[1] = c % A.cdim[1];
        coord[2] = c / A.cdim[1];
        const int limb = (int)coord[A.limb_axis];
        const LimbInfo& L = A.limbs[limb];
        const ModInfo mi = limb_modinfo(L);
        const uint64_t q = L.q;
        const int64_t so = col_offset(A.src, coord), dof = col_offset(A.dst, coord);
        const int64_t oo = A.old.ptr ? col_offset(A.old, coord) : 0;
        // the network's input: a, scattered by perm for the inverse
        // (x[perm[u]] = a[u]; rows hit by no entry keep the result's contents)
        for (int64_t j = threadIdx.x; j < m; j += blockDim.x) {
            int64_t v;
            if (A.inverse && A.perm_kind == PERM_MAP) {
                v = load_any(A.src, so + (int64_t)A.rowmap[j] * A.src.row_stride);
            } else if (A.inverse && A.perm_kind == PERM_SCATTER_INV) {
                const int32_t r = A.rowmap[j];
                v = r >= 0 ? load_any(A.src, so + (int64_t)r * A.src.row_stride)
                           : load_any(A.old, oo + j * A.old.row_stride);
            } else {
                v = load_any(A.src, so + j * A.src.row_stride);
            }
            x[j] = v;
        }
        __syncthreads();
        const int64_t* twr = A.raw_tw ? A.raw_tw + (int64_t)limb * m : nullptr;
        const uint64_t* twp = A.tw + 2 * (int64_t)limb * m;
        if (!A.inverse) {
            // stage s = 2^lg: butterflies (jx, jx + step), step = m / 2s, twiddle tw[t, s + u]
            for (int lg = 0; lg < A.logm; ++lg) {
                const int sb = A.logm - 1 - lg;
                const int64_t stage = int64_t(1) << lg, step = int64_t(1) << sb;
                for (int64_t b = threadIdx.x; b < half; b += blockDim.x) {
                    const int64_t u = b >> sb, jx = (u << (sb + 1)) + (b & (step - 1));
                    const int64_t s = twr ? twr[stage + u] : (int64_t)twp[2 * (stage + u)];
                    const int64_t uv = x[jx];
                    const int64_t vm = ref_mulmod(mi, A.is32, x[jx + step], s);
                    x[jx] = ref_t(cpp_rem(mi, (int64_t)((uint64_t)uv + (uint64_t)vm)), A.is32);
                    x[jx + step] = ref_t(cpp_rem(mi, (int64_t)((uint64_t)uv + q - (uint64_t)vm)), A.is32);
                }
                __syncthreads();
            }
            // out[u] = x[perm[u]]
            for (int64_t j = threadIdx.x; j < m; j += blockDim.x) {
                int64_t row = j, v = x[j];
                if (A.perm_kind == PERM_MAP) row = A.rowmap[j];
                else if (A.perm_kind == PERM_GATHER_FWD) v = x[A.rowmap[j]];
                store_any(A.dst, dof + row * A.dst.row_stride, (uint64_t)v);
            }
        } else {
            // half = m/2 .. 1 with t_stride = 1 .. m/2, twiddle itw[t, half + group]
            for (int lg = 0; lg < A.logm; ++lg) {
                const int64_t ts = int64_t(1) << lg, hf = half >> lg;
                for (int64_t tid = threadIdx.x; tid < half; tid += blockDim.x) {
                    const int64_t group = tid >> lg;
                    const int64_t iu = (group << (lg + 1)) + (tid & (ts - 1)), iv = iu + ts;
                    const int64_t s = twr ? twr[hf + group] : (int64_t)twp[2 * (hf + group)];
                    const int64_t uv = x[iu], vv = x[iv];
                    x[iu] = ref_t(cpp_rem(mi, (int64_t)((uint64_t)uv + (uint64_t)vv)), A.is32);
                    x[iv] = ref_mulmod(mi, A.is32, ref_t((int64_t)((uint64_t)uv + q - (uint64_t)vv), A.is32), s);
                }
                __syncthreads();
            }
            const int64_t mv = A.raw_minv ? A.raw_minv[limb] : (int64_t)L.minv;
            for (int64_t j = threadIdx.x; j < m; j += blockDim.x)
                store_any(A.dst, dof + j * A.dst.row_stride, (uint64_t)ref_mulmod(mi, A.is32, x[j], mv));
        }
        __syncthreads();
        if (threadIdx.x == 0) A.colflag[col] = 0u;
    }
}

// raw copies of the tables for ref_all plans: out[t][i] = tw[t, i], minv_out[t] = m_inv[t]
__global__ void k_plan_raw(PlanSrc tw, PlanSrc minv, int has_minv, int64_t k, int64_t m, int64_t* out,
                           int64_t* minv_out) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < k * m; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = e / m, i = e - t * m;
        out[e] = ld_elem(tw.ptr, t * tw.s0 + i * tw.s1, tw.elem32);
        if (has_minv && i == 0) minv_out[t] = ld_elem(minv.ptr, t * minv.s0, minv.elem32);
    }
}

// ------------------------------------------------------------ host side
int ilog2_exact(int64_t m) {
    int l = 0;
    while ((int64_t(1) << l) < m) ++l;
    return ((int64_t(1) << l) == m) ? l : -1;
}

void require_int_view(const hx_view* v, const char* name) {
    check_view(v, name);
    if (v->dtype != HX_INT32 && v->dtype != HX_INT64)
        raise(HX_ERR_INVALID_ARGUMENT, std::string(name) + ": NTT needs int32 or int64 tensors");
}

PlanSrc plan_src(const hx_view* v) {
    PlanSrc s;
    s.ptr = v->data;
    s.s0 = v->ndim > 0 ? v->strides[0] : 0;
    s.s1 = v->ndim > 1 ? v->strides[1] : 0;
    s.elem32 = v->dtype == HX_INT32;
    return s;
}



ColAddr tensor_addr(const hx_view* v, const int* order) {
    ColAddr c;
    c.ptr = v->data;
    c.row_stride = v->strides[1];
    const int axes[3] = {0, 2, 3};
    for (int i = 0; i < 3; ++i) c.cstride[i] = v->strides[axes[order[i]]];
    c.kind = v->dtype == HX_INT32 ? 1 : 0;
    return c;
}

// intermediate with the rows of a column contiguous: column c at ptr + c * m
ColAddr tmp_addr_rows(void* ptr, int64_t m) {
    ColAddr c;
    c.ptr = ptr;
    c.row_stride = 1;
    c.cstride[0] = m;
    c.cstride[1] = 0;
    c.cstride[2] = 0;
    c.kind = 2;
    return c;
}

ColAddr tmp_addr(void* ptr, int64_t ncols) {
    ColAddr c;
    c.ptr = ptr;
    c.row_stride = ncols;
    c.cstride[0] = 1;
    c.cstride[1] = 0;
    c.cstride[2] = 0;
    c.kind = 2;
    return c;
}

bool r16_window(int H) {
    static const bool generic_only = [] {
        const char* e = getenv("HEAL_AMD_NTT_GENERIC");
        return e && e[0] == '1';
    }();
    return !generic_only && H >= 6 && H <= 8;
}

constexpr int kR16Tile = 32;  // sub-problems per workgroup of the radix-16 pass

// Launches one pass.  P.S_cols / P.S_bot give the tile (sub-problems per
// workgroup, see configure_tile); a radix-16 sub-problem takes 2^(H-4) lanes.
// Fills P.n_bot_tiles / P.n_col_tiles and returns the pass's workgroup count.
int64_t pass_blocks(PassArgs& P) {
    const int64_t ntop = int64_t(1) << (P.logm - P.b0 - P.H);
    if (P.top_bitrev) {  // one column and 32 tops per tile
        const int64_t blocks = P.ncols * (ntop / kR16Tile);
        if (blocks > 0x7fffffff) raise(HX_ERR_INVALID_ARGUMENT, "NTT too large for one launch");
        return blocks;
    }
    P.n_bot_tiles = (P.bot_count + P.S_bot - 1) / P.S_bot;
    P.n_col_tiles = (P.ncols + P.S_cols - 1) / P.S_cols;
    const int64_t blocks = ntop * P.n_bot_tiles * P.n_col_tiles;
    if (blocks > 0x7fffffff || P.ncols > 0x7fffffff) raise(HX_ERR_INVALID_ARGUMENT, "NTT too large for one launch");
    P.fd_ctiles = make_fastdiv((uint32_t)std::max<int64_t>(1, P.n_col_tiles));
    P.fd_btiles = make_fastdiv((uint32_t)std::max<int64_t>(1, P.n_bot_tiles));
    P.fd_c0 = make_fastdiv((uint32_t)std::max<int64_t>(1, P.cdim[0]));
    P.fd_c1 = make_fastdiv((uint32_t)std::max<int64_t>(1, P.cdim[1]));
    P.log_scols = 0;
    while ((1 << P.log_scols) < P.S_cols) ++P.log_scols;
    return blocks;
}

void launch_pass(PassArgs& P, hipStream_t st) {
    const int W = 1 << P.H;
    const int S = P.S_cols * P.S_bot;
    const int64_t blocks = pass_blocks(P);
    if (P.rm.on && (P.top_bitrev || !r16_window(P.H)))
        raise(HX_ERR_RUNTIME, "internal: accumulate epilogue needs a radix-16 pass without bit-reversal tiles");
    if (P.top_bitrev) {
        if (!r16_window(P.H) || P.b0 != 0 || P.logm - P.H < 5) raise(HX_ERR_RUNTIME, "internal: bit-reversal tiles");
        const size_t lds = (size_t)kR16Tile * (W + 1) * 8;
        const dim3 grid((unsigned)blocks), block((unsigned)(kR16Tile * (W / 16)));
#define HX_R16B(HH)                                                                                  \
    if (P.inverse) hipLaunchKernelGGL((k_ntt_pass_r16_brt<HH, true>), grid, block, lds, st, P);      \
    else hipLaunchKernelGGL((k_ntt_pass_r16_brt<HH, false>), grid, block, lds, st, P);
        if (P.H == 8) {
            HX_R16B(8)
        } else if (P.H == 7) {
            HX_R16B(7)
        } else {
            HX_R16B(6)
        }
#undef HX_R16B
        check_hip(hipGetLastError(), "ntt pass launch");
        return;
    }
    if (r16_window(P.H)) {
        const bool lm = P.lane_major != 0;
        if (S != kR16Tile && !(S == 16 && !lm)) raise(HX_ERR_RUNTIME, "internal: radix-16 tile size");
        const size_t lds = (size_t)S * (lm ? W + W / 16 : W + 1) * 8;
        const dim3 grid((unsigned)blocks), block((unsigned)(S * (W / 16)));
#define HX_R16(HH, LMV)                                                                                     \
    if (P.inverse) hipLaunchKernelGGL((k_ntt_pass_r16<HH, true, LMV, false, 32>), grid, block, lds, st, P); \
    else hipLaunchKernelGGL((k_ntt_pass_r16<HH, false, LMV, false, 32>), grid, block, lds, st, P);
#define HX_R16S(HH, RMV, SV)                                                                                \
    if (P.inverse) hipLaunchKernelGGL((k_ntt_pass_r16<HH, true, false, RMV, SV>), grid, block, lds, st, P); \
    else hipLaunchKernelGGL((k_ntt_pass_r16<HH, false, false, false, SV>), grid, block, lds, st, P);
        if (P.rm.on && (!P.inverse || lm))
            raise(HX_ERR_RUNTIME, "internal: accumulate epilogue on a forward or lane-major pass");
        if (P.rm.on && S == 32) {
            if (P.H == 8) {
                HX_R16S(8, true, 32)
            } else if (P.H == 7) {
                HX_R16S(7, true, 32)
            } else {
                HX_R16S(6, true, 32)
            }
        } else if (P.rm.on) {
            if (P.H == 8) {
                HX_R16S(8, true, 16)
            } else if (P.H == 7) {
                HX_R16S(7, true, 16)
            } else {
                HX_R16S(6, true, 16)
            }
        } else if (S == 16) {
            if (P.H == 8) {
                HX_R16S(8, false, 16)
            } else if (P.H == 7) {
                HX_R16S(7, false, 16)
            } else {
                HX_R16S(6, false, 16)
            }
        } else if (P.H == 8) {
            if (lm) {
                HX_R16(8, true)
            } else {
                HX_R16(8, false)
            }
        } else if (P.H == 7) {
            if (lm) {
                HX_R16(7, true)
            } else {
                HX_R16(7, false)
            }
        } else {
            if (lm) {
                HX_R16(6, true)
            } else {
                HX_R16(6, false)
            }
        }
#undef HX_R16S
#undef HX_R16
    } else {
        const size_t lds = (size_t)S * (W + 1) * 8;
        hipLaunchKernelGGL(k_ntt_pass, dim3((unsigned)blocks), dim3(kThreads), lds, st, P);
    }
    check_hip(hipGetLastError(), "ntt pass launch");
}

// Tile shape of a pass: col_tiles puts adjacent columns of one row in a tile
// (layouts with the limb axis innermost), otherwise adjacent rows of one
// column (limb-major layouts).  P.H, P.ncols and P.bot_count must be set.
// r16_tile: sub-problems per workgroup of a radix-16 pass, 32 or (a tile of
// adjacent columns only) 16.
void configure_tile(PassArgs& P, bool col_tiles, int npass, int r16_tile = kR16Tile) {
    if (r16_window(P.H)) {
        const int T = col_tiles && r16_tile == 16 ? 16 : kR16Tile;
        if (col_tiles) {
            int sc = 1;
            while (sc < T && sc < P.ncols) sc <<= 1;
            P.S_cols = sc;
            P.S_bot = T / sc;
        } else {
            P.S_cols = 1;
            P.S_bot = T;
        }
        return;
    }
    const int W = 1 << P.H;
    int S = W >= 256 ? 8 : std::max(1, std::min(64, 2048 / W));
    if (npass == 1 && W > 256) S = std::max(1, 4096 / W);
    if (!col_tiles) {
        P.S_cols = 1;
        P.S_bot = (int)std::max<int64_t>(1, std::min<int64_t>(S, P.bot_count));
    } else {
        P.S_cols = S;
        P.S_bot = 1;
    }
}

}  // namespace
}  // namespace hx
