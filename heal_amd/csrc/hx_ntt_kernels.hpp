// hx_ntt_kernels.hpp — NTT device code shared by hx_ntt.hip and hx_keyswitch.hip:
// plan layout, pass arguments, the generic and register-radix-8 pass kernels,
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
    uint64_t pad[2];
};

struct hx_ntt_plan {
    int device;
    int inverse;
    int dtype;
    int logm;
    int64_t m;
    int64_t k;
    int perm_kind;        // PERM_* below
    void* mem;            // single device allocation
    // Intermediate buffer between passes.  A plan executes on one stream at a
    // time (like an FFT plan); the workspace grows on first use.
    mutable void* ws;
    mutable size_t ws_bytes;
    LimbInfo* limbs;      // [k]
    uint64_t* tw;         // [k][m][2]: (w mod q, companion)
    int32_t* rowmap;      // [m] (see perm_kind)
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

__global__ void k_plan_limbs(PlanSrc p, PlanSrc minv, int has_minv, int64_t k, LimbInfo* limbs,
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
    ModInfo mi = make_modinfo((uint64_t)q);
    L.q = (uint64_t)q;
    L.d = mi.d;
    L.v = mi.v;
    L.s = mi.s;
    L.small = q < (int64_t(1) << 31);
    if (has_minv) {
        uint64_t mv = residue(mi, ld_elem(minv.ptr, t * minv.s0, minv.elem32));
        L.minv = mv;
        L.minv_p = L.small ? divq128(mi, 0, mv << 32) : shoup_companion(mi, mv);
    }
    limbs[t] = L;
}

__global__ void k_plan_twiddles(PlanSrc tw, int64_t k, int64_t m, const LimbInfo* limbs, uint64_t* out) {
    int64_t n = k * m;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t t = e / m, i = e - t * m;
        LimbInfo L = limbs[t];
        ModInfo mi = limb_modinfo(L);
        uint64_t w = residue(mi, ld_elem(tw.ptr, t * tw.s0 + i * tw.s1, tw.elem32));
        uint64_t wp = L.small ? divq128(mi, 0, w << 32) : shoup_companion(mi, w);
        out[2 * e] = w;
        out[2 * e + 1] = wp;
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
    int32_t reduce_in;
    int32_t scale_out;
    int32_t lazy_out;        // radix-8 passes: leave 64-bit residues in [0, 4q) for the next radix-8 pass
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
};

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

// Per-tile bookkeeping shared by the pass kernels (one entry per sub-problem).
struct TileShared {
    int64_t src_off[64];
    int64_t dst_off[64];
    int64_t alt_off[64];
    int64_t jbase[64];
    int32_t limb[64];
    int32_t digit[64];
    int32_t valid[64];
};

__device__ __forceinline__ void tile_setup(const PassArgs& P, int S, TileShared& T) {
    int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t col_tile = tile % P.n_col_tiles;
    tile /= P.n_col_tiles;
    const int64_t bot_tile = tile % P.n_bot_tiles;
    const int64_t top = tile / P.n_bot_tiles;
    if (threadIdx.x < S) {
        int s = threadIdx.x;
        int cs = s % P.S_cols, bs = s / P.S_cols;
        int64_t col = col_tile * P.S_cols + cs;
        int64_t bot = bot_tile * P.S_bot + bs;
        int valid = col < P.ncols && bot < P.bot_count;
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
    }
    __syncthreads();
}

// Coalesced load of the tile into LDS: consecutive threads take consecutive
// sub-problems of one window row.  `slot(s, w)` gives the LDS index.
template <typename Slot>
__device__ __forceinline__ void tile_load(const PassArgs& P, int S, int W, const TileShared& T, uint64_t* xs,
                                          Slot slot) {
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
    tile_setup(P, S, T);
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
// Register-radix-8 pass for windows of H = 6..8 bits.
//
// A sub-problem (2^H points) is held by 2^(H-3) lanes with 8 values each.  The
// window's stages run in three rounds of at most 3 stages (bits [H-3, H),
// [H-6, H-3), [0, H-6)); inside a round every butterfly is between two
// registers of one lane, and the rounds exchange values through LDS.  In a
// round over bits [r0, r0 + nb), value e of lane `lane` is window index
//     w = (rest >> r0) << (r0 + nb) | (e mod 2^nb) << r0 | (rest mod 2^r0),
//     rest = (e >> nb) << (H - 3) | lane.
// LDS slots are padded by one every 32 entries to spread banks.
//
// Arithmetic per limb (uniform over a sub-problem):
//   MODE_S32  q < 2^31: 32-bit Shoup, residues kept in [0, q);
//   MODE_L64  q < 2^62: 64-bit Shoup with Harvey's lazy butterflies
//             (D. Harvey, "Faster arithmetic for number-theoretic transforms",
//             J. Symb. Comp. 60, 2014): forward values stay in [0, 4q),
//             inverse values in [0, 2q), reduced when leaving the transform;
//   MODE_F64  q < 2^63: 64-bit Shoup, residues kept in [0, q).
enum : int { MODE_S32 = 0, MODE_L64 = 1, MODE_F64 = 2 };

__device__ __forceinline__ int limb_mode(const LimbInfo& L) {
    return L.small ? MODE_S32 : (L.q < (uint64_t(1) << 62) ? MODE_L64 : MODE_F64);
}

template <int H>
__device__ __forceinline__ int r8_w(int lane, int e, int r0, int nb) {
    int v = e & ((1 << nb) - 1);
    int rest = ((e >> nb) << (H - 3)) | lane;
    return ((rest >> r0) << (r0 + nb)) | (v << r0) | (rest & ((1 << r0) - 1));
}

__device__ __forceinline__ int r8_pad(int w) { return w + (w >> 5); }

struct TwRow {
    const ulonglong2* tw;  // (w, companion) pairs of this limb
    int64_t top;
    int logm, g0;
};

template <int MODE, bool INV>
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
        const uint64_t q2 = q << 1;
        if (!INV) {  // x0, x1 in [0, 4q) -> [0, 4q)
            uint64_t u = x0 >= q2 ? x0 - q2 : x0;
            uint64_t t = shoup_lazy(x1, w, wp, q);  // [0, 2q)
            x0 = u + t;
            x1 = u - t + q2;
        } else {  // x0, x1 in [0, 2q) -> [0, 2q)
            uint64_t a = x0 + x1;
            uint64_t d = x0 - x1 + q2;
            x0 = a >= q2 ? a - q2 : a;
            x1 = shoup_lazy(d, w, wp, q);
        }
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

// One round (nb stages) on the 8 registers of a lane.  The twiddle index of a
// butterfly depends on the bits of its lower element above the stage's window
// bit; writing it as (rest_high, round bits above bi) lets equal indices
// within a lane fold into one load.
template <int H, bool INV, int MODE>
__device__ __forceinline__ void r8_round(uint64_t (&x)[8], int lane, int r0, int nb, const TwRow& R, uint64_t q) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= nb) break;
        const int bi = INV ? k : (nb - 1 - k);  // bit of e
        const int wbit = r0 + bi;               // window bit
        const int st = INV ? wbit : (H - 1 - wbit);
        const int g = R.g0 + st;
        const int64_t base = INV ? ((int64_t(1) << (R.logm - 1 - g)) + (R.top << (H - 1 - st)))
                                 : ((int64_t(1) << g) + (R.top << st));
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            if (e & (1 << bi)) continue;
            const int e2 = e | (1 << bi);
            // (w_lo >> (wbit + 1)) = rest_high << (nb - bi - 1) | (v >> (bi + 1))
            const int rest_high = (((e >> nb) << (H - 3)) | lane) >> r0;
            const int hi = (rest_high << (nb - bi - 1)) | ((e & ((1 << nb) - 1)) >> (bi + 1));
            ulonglong2 tw = R.tw[base + hi];
            butterfly<MODE, INV>(x[e], x[e2], tw.x, tw.y, q);
        }
    }
}

template <int H, bool INV, int MODE>
__device__ __forceinline__ void r8_round_lds(uint64_t* xs_s, int lane, int r0, int nb, const TwRow& R, uint64_t q) {
    uint64_t x[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) x[e] = xs_s[r8_pad(r8_w<H>(lane, e, r0, nb))];
    r8_round<H, INV, MODE>(x, lane, r0, nb, R, q);
#pragma unroll
    for (int e = 0; e < 8; ++e) xs_s[r8_pad(r8_w<H>(lane, e, r0, nb))] = x[e];
}

// All rounds of one sub-problem.  Every thread of the workgroup reaches the
// same barriers; only the arithmetic between them depends on the limb mode.
// Threads with active == false (no sub-problem of their own) only take part
// in the barriers.
template <int H, bool INV>
__device__ __forceinline__ void r8_all_rounds(uint64_t* xs_s, int lane, const TwRow& R, uint64_t q, int mode,
                                              bool active = true) {
    constexpr int nC = H - 6;
    // forward: rounds A (top bits), B, C; inverse: C, B, A
#pragma unroll
    for (int ri = 0; ri < 3; ++ri) {
        const int r = INV ? 2 - ri : ri;
        const int r0 = r == 0 ? H - 3 : (r == 1 ? H - 6 : 0);
        const int nb = r == 2 ? nC : 3;
        if (nb == 0) continue;
        if (active) {
            if (mode == MODE_L64) r8_round_lds<H, INV, MODE_L64>(xs_s, lane, r0, nb, R, q);
            else if (mode == MODE_S32) r8_round_lds<H, INV, MODE_S32>(xs_s, lane, r0, nb, R, q);
            else r8_round_lds<H, INV, MODE_F64>(xs_s, lane, r0, nb, R, q);
        }
        __syncthreads();
    }
}

// Fully reduce a lazy value of a MODE_L64 limb.
__device__ __forceinline__ uint64_t lazy_reduce(uint64_t x, uint64_t q) {
    const uint64_t q2 = q << 1;
    x = x >= q2 ? x - q2 : x;
    return x >= q ? x - q : x;
}

template <int H, bool INV, int THREADS>
__global__ __launch_bounds__(THREADS) void k_ntt_pass_r8(PassArgs P) {
    constexpr int W = 1 << H;
    constexpr int WP = W + W / 32;
    constexpr int TL = W / 8;          // lanes per sub-problem
    constexpr int S = THREADS / TL;    // sub-problems per workgroup
    extern __shared__ uint64_t xs[];   // [S][WP]
    __shared__ TileShared T;
    tile_setup(P, S, T);

    // ---- load: thread t always serves sub-problem t % S, rows t / S + k * TL
    {
        const int s = threadIdx.x % S;
        const int w0 = threadIdx.x / S;
        if (T.valid[s]) {
            const LimbInfo& L = P.limbs[T.limb[s]];
            const uint64_t q = L.q;
            uint64_t* xs_s = xs + s * WP;
            const int es = P.src.kind == 1 ? 4 : 8;
            if (P.src.kind == 3) {
                // digit d of x: the reference's unsigned digit loop
                // (example_impl/g_decomposition_impl.cpp:145-149) in closed form
                const char* base = (const char*)P.src.ptr + 8 * (T.src_off[s] + T.jbase[s] * P.src.row_stride);
                const int64_t wstride = (int64_t)8 * (P.src.row_stride << P.b0);
                const int d = T.digit[s];
                const int wb = P.gd_w;
                const uint64_t mask = wb ? (~0ull >> (64 - wb)) : 0;
                const int shift = wb * d;
                const int rest = wb * P.gd_D;
                unsigned long long over = 0;
                for (int w = w0; w < W; w += TL) {
                    const int64_t x = *(const int64_t*)(base + (int64_t)w * wstride);
                    const uint64_t u = (uint64_t)x;
                    uint64_t v = (wb && shift < 64) ? ((u >> shift) & mask) : 0;
                    if (d == 0) over += wb ? (rest < 64 && (u >> rest) != 0) : (x > 0);
                    if (P.reduce_in && v >= q) v = residue(limb_modinfo(L), (int64_t)v);
                    xs_s[r8_pad(w)] = v;
                }
                if (over) atomicAdd(P.gd_status, over);
            } else if (!P.src_map) {
                const char* base = (const char*)P.src.ptr + es * (T.src_off[s] + T.jbase[s] * P.src.row_stride);
                const int64_t wstride = (int64_t)es * (P.src.row_stride << P.b0);
#pragma unroll 4
                for (int w = w0; w < W; w += TL) {
                    const char* ptr = base + (int64_t)w * wstride;
                    int64_t x = es == 4 ? (int64_t) * (const int32_t*)ptr : *(const int64_t*)ptr;
                    uint64_t v = (uint64_t)x;
                    if (P.reduce_in && (x < 0 || v >= q)) v = residue(limb_modinfo(L), x);
                    xs_s[r8_pad(w)] = v;
                }
            } else {
                for (int w = w0; w < W; w += TL) {
                    int64_t j = T.jbase[s] + ((int64_t)w << P.b0);
                    int32_t r = P.src_map[j];
                    int64_t x = r >= 0 ? load_any(P.src, T.src_off[s] + (int64_t)r * P.src.row_stride)
                                       : load_any(P.alt, T.alt_off[s] + j * P.alt.row_stride);
                    uint64_t v = (uint64_t)x;
                    if (P.reduce_in && (x < 0 || v >= q)) v = residue(limb_modinfo(L), x);
                    xs_s[r8_pad(w)] = v;
                }
            }
        }
    }
    __syncthreads();

    // ---- the window's stages
    {
        const int s = threadIdx.x / TL, lane = threadIdx.x % TL;
        const LimbInfo& L = P.limbs[T.limb[s]];
        TwRow R;
        R.tw = (const ulonglong2*)(P.tw + 2 * (int64_t)T.limb[s] * P.m);
        R.top = T.jbase[s] >> (P.b0 + H);
        R.logm = P.logm;
        R.g0 = P.stage0;
        // invalid sub-problems (past the last column) compute on stale LDS
        // values and are never stored
        r8_all_rounds<H, INV>(xs + s * WP, lane, R, L.q, limb_mode(L));
    }

    // ---- store
    {
        const int s = threadIdx.x % S;
        const int w0 = threadIdx.x / S;
        if (!T.valid[s]) return;
        const LimbInfo& L = P.limbs[T.limb[s]];
        const uint64_t q = L.q;
        const int mode = limb_mode(L);
        const uint64_t* xs_s = xs + s * WP;
        const int es = P.dst.kind == 1 ? 4 : 8;
        const bool full = !P.lazy_out;
        const int64_t colbase = T.dst_off[s];
        for (int w = w0; w < W; w += TL) {
            uint64_t v = xs_s[r8_pad(w)];
            if (P.scale_out) {
                if (mode == MODE_S32) v = shoup_mul32((uint32_t)v, (uint32_t)L.minv, (uint32_t)L.minv_p, (uint32_t)q);
                else v = shoup_mul(v, L.minv, L.minv_p, q);
            } else if (full && mode == MODE_L64) {
                v = lazy_reduce(v, q);
            }
            int64_t j = T.jbase[s] + ((int64_t)w << P.b0);
            int64_t row = P.dst_map_limb ? (int64_t)P.dst_map_limb[(int64_t)T.limb[s] * P.m + j]
                                         : (P.dst_map ? (int64_t)P.dst_map[j] : j);
            char* ptr = (char*)P.dst.ptr + es * (colbase + row * P.dst.row_stride);
            if (es == 4) *(int32_t*)ptr = (int32_t)v;
            else *(uint64_t*)ptr = v;
        }
    }
}

// dst[u][col] = src[map[u]][col] (forward gather for non-permutation perms)
struct GatherArgs {
    ColAddr src;
    ColAddr dst;
    int64_t ncols;
    int64_t cdim[3];
    const int32_t* map;
    int64_t m;
};

__global__ void k_row_gather(GatherArgs G) {
    int64_t n = G.m * G.ncols;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t u = e / G.ncols, col = e - u * G.ncols;
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

bool r8_window(int H) {
    static const bool generic_only = [] {
        const char* e = getenv("HEAL_AMD_NTT_GENERIC");
        return e && e[0] == '1';
    }();
    return !generic_only && H >= 6 && H <= 8;
}

// Launches one pass.  P.S_cols / P.S_bot give the tile (sub-problems per
// workgroup); for register-radix-8 windows each sub-problem takes 2^(H-3) lanes.
void launch_pass(PassArgs& P, hipStream_t st) {
    const int W = 1 << P.H;
    int S = P.S_cols * P.S_bot;
    int64_t ntop = int64_t(1) << (P.logm - P.b0 - P.H);
    P.n_bot_tiles = (P.bot_count + P.S_bot - 1) / P.S_bot;
    P.n_col_tiles = (P.ncols + P.S_cols - 1) / P.S_cols;
    int64_t blocks = ntop * P.n_bot_tiles * P.n_col_tiles;
    if (blocks > 0x7fffffff) raise(HX_ERR_INVALID_ARGUMENT, "NTT too large for one launch");
    if (r8_window(P.H)) {
        const int TL = W / 8;
        const int threads = S * TL;
        if (threads != 256 && threads != 512) raise(HX_ERR_RUNTIME, "internal: radix-8 tile size");
        size_t lds = (size_t)S * (W + W / 32) * 8;
        dim3 grid((unsigned)blocks), block((unsigned)threads);
#define HX_R8(HH, TT)                                                                              \
    if (P.inverse) hipLaunchKernelGGL((k_ntt_pass_r8<HH, true, TT>), grid, block, lds, st, P);     \
    else hipLaunchKernelGGL((k_ntt_pass_r8<HH, false, TT>), grid, block, lds, st, P);
        if (threads == 512) {
            if (P.H == 8) {
                HX_R8(8, 512)
            } else if (P.H == 7) {
                HX_R8(7, 512)
            } else {
                HX_R8(6, 512)
            }
        } else {
            if (P.H == 8) {
                HX_R8(8, 256)
            } else if (P.H == 7) {
                HX_R8(7, 256)
            } else {
                HX_R8(6, 256)
            }
        }
#undef HX_R8
    } else {
        size_t lds = (size_t)S * (W + 1) * 8;
        hipLaunchKernelGGL(k_ntt_pass, dim3((unsigned)blocks), dim3(kThreads), lds, st, P);
    }
    check_hip(hipGetLastError(), "ntt pass launch");
}

// Sub-problems per workgroup for a window of H bits.
int tile_subproblems(int H, int npass, bool many_cols) {
    const int W = 1 << H;
    if (r8_window(H)) {
        const int TL = W / 8;
        return (many_cols ? 512 : 256) / TL;
    }
    int S = W >= 256 ? 8 : std::max(1, std::min(64, 2048 / W));
    if (npass == 1 && W > 256) S = std::max(1, 4096 / W);
    return S;
}

}  // namespace
}  // namespace hx
