// hx_keyswitch.hip — fused hybrid key-switch (vendor extension of the HEAL API).
//
// Computes, for B independent instances, exactly what the six HEAL calls
//     g_decomposition -> ntt -> modmul_ttt -> axis_modsum(axis 0) -> intt -> permute
// compute on the views that chain uses (x [B, N, L] -> digits [B, N, L, D]
// -> digit NTTs [B, N, D, L] -> products with the evaluation key
// [D, B, 2, N, L] -> sum over digits [B, 2, N, L] -> INTT -> Galois gather with
// perms [L, N]), in three kernels:
//   K1  forward pass over the high index bits, reading x and extracting the
//       base-2^w digits on load;
//   K2  forward pass over the low index bits of all D digit columns of a limb
//       group, the key inner product and digit sum in LDS, and the inverse
//       pass over the low bits;
//   K3  inverse pass over the high bits, m^-1 scaling, and the Galois
//       permutation applied as a row scatter on store (one shared row map
//       when every limb has the same permutation).
// Between the forward and inverse transforms the values stay in the NTT's
// network order; the key is permuted into that order once, as residues
// (hx_keyswitch_prepare_key), which is why the forward and inverse plans must
// share one permutation.  The digits follow the reference's unsigned digit
// loop and every other step is exact modular arithmetic on the residues the
// chain's INTT receives, so the result is bit-identical with the six-call chain.
#include <sys/stat.h>

#include <algorithm>
#include <cstdio>
#include <map>
#include <mutex>
#include <new>
#include <tuple>
#include <vector>

#include "hx_ntt_kernels.hpp"

struct hx_keyswitch_plan {
    int device;
    const hx_ntt_plan* fwd;
    const hx_ntt_plan* inv;
    int64_t m, k;
    int logm, hhi, hlo;
    int D, w;
    int G;            // limbs per K2 workgroup (tuned)
    int tile;         // sub-problems per K1/K3 workgroup, 32 or 16 (tuned)
    void* mem;
    int32_t* galmap;  // [k][m]: output row of coefficient j of limb t
    int gal_shared;   // every limb has the same map (K3 then reads row 0 only)
    // K1 -> K2 and K2 -> K3 intermediates, per stream (see hx::Workspaces)
    hx::Workspaces* wss;
};

namespace hx {
namespace {

struct InnerArgs {
    const uint64_t* src;  // K1 output: rows of D*k*B, column d + D*(t + k*b)
    uint64_t* dst;        // K3 input:  rows of 2*k*B, column c + 2*(t + k*b)
    const int64_t* key;   // [D][key_b][2][key_k][m], window-local order (k_ks_prepare_key)
    int64_t key_k;        // limbs the key was prepared with (>= k; the first k are used)
    int64_t key_b;        // keys in the buffer: 1 = one key for every instance, else B
    const LimbInfo* limbs;
    const uint64_t* tw_f;
    const uint64_t* tw_i;
    int64_t m, k, B;
    int32_t logm, hhi, D, G;
    int64_t n_tg, n_hi;
    int32_t H;            // the inner window (= logm - hhi)
    int32_t tpt, per;     // grouped launches: threads per tile, tiles per workgroup
    FastDiv fd_B, fd_ntg; // 32-bit decoding of the tile index (divisors B and n_tg)
};

// LDS rows of K2 (64-bit slots): one pad slot every 16 and one more per row,
// W + W/16 + 1 slots.  A 64-bit access hits bank pair (slot mod 16).  Threads
// with adjacent lanes of one sub-problem spread over the bank pairs through
// the pad slots; threads with adjacent sub-problems and one lane through the
// row length, which is 1 mod 16, so that the exchanges with sub-problems on
// adjacent threads see at most 2-way conflicts (a row length of 0 mod 16 put
// every sub-problem of a lane on one bank pair).
__host__ __device__ constexpr int ks_lds_row_of(int H) {
    return (1 << H) + (1 << H) / 16 + 1;
}
template <int H>
__device__ constexpr int ks_lds_row() {
    return ks_lds_row_of(H);
}
__device__ __forceinline__ int ks_slot(int w) {
    return w + (w >> 4);
}

// The forward values of a MODE_L64 limb ([0, 16q)) enter the key inner product
// partly reduced, in [0, 4q), when the sum of D products stays below q * 2^64:
// D * 4q * q <= q * 2^64.
__device__ __forceinline__ bool ks_lazy_dot(const LimbInfo& L, int D) {
    return limb_mode(L) == MODE_L64 && (uint64_t)(4 * D) <= ~0ull / L.q;
}

// Limbs whose prepared key is in Montgomery form, key * 2^64 mod q (odd q of
// the lazy 64-bit class): the inner product then ends with one Montgomery
// reduction, whose result in [0, 2q) the lazy inverse butterflies take as is.
__device__ __forceinline__ bool ks_mont(const LimbInfo& L) {
    return limb_mode(L) == MODE_L64 && L.qni != 0;
}

// Key inner product of one lane's 16 values of layout B:
//     x[e] = (sum_d F_d[w] * key_d[e]) mod q,   w = r16_wB(e, lane),
// F_d the forward values of digit d in LDS (rows SP apart), key words in
// [0, q).  One 128-bit accumulation and one reduction when the sum stays below
// q * 2^64 (`single`), else a reduction per digit.  MONT: the key words are
// key_d * 2^64 mod q and the reduction is Montgomery's with qni = -q^-1 mod 2^64
// (result in [0, 2q)); otherwise Möller–Granlund (result in [0, q)).  DC > 0: D = DC at compile
// time, so that all key loads of a lane can be in flight together.
template <int H, int DC, bool MONT>
__device__ __forceinline__ void ks_dot(uint64_t (&x)[16], const uint64_t* xf, const int64_t* kp, int64_t dstride,
                                       int lane, const ModInfo& mi, int D, bool single, uint64_t qni) {
    constexpr int SP = ks_lds_row<H>();
    constexpr int TL = (1 << H) / 16;
    if (DC > 0) D = DC;
    if (DC > 0 && single) {
        // the key words of CH registers are requested together, then used
        constexpr int CH = DC <= 3 ? 16 : 8;
#pragma unroll
        for (int e0 = 0; e0 < 16; e0 += CH) {
            uint64_t kv[DC > 0 ? DC : 1][CH];
#pragma unroll
            for (int d = 0; d < DC; ++d)
#pragma unroll
                for (int i = 0; i < CH; ++i) kv[d][i] = (uint64_t)kp[d * dstride + (e0 + i) * TL];
#pragma unroll
            for (int i = 0; i < CH; ++i) {
                const int w = r16_wB<H>(e0 + i, lane);
                u128 acc = 0;
#pragma unroll
                for (int d = 0; d < DC; ++d) acc += (u128)xf[d * SP + ks_slot(w)] * kv[d][i];
                x[e0 + i] = MONT ? redc((uint64_t)(acc >> 64), (uint64_t)acc, mi.M, qni)
                                 : reduce128(mi, (uint64_t)(acc >> 64), (uint64_t)acc);
            }
        }
    } else if (single) {
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int w = r16_wB<H>(e, lane);
            u128 acc = 0;
#pragma unroll
            for (int d = 0; d < (DC > 0 ? DC : 1); ++d)
                acc += (u128)xf[d * SP + ks_slot(w)] * (uint64_t)kp[d * dstride + e * TL];
            for (int d = DC > 0 ? DC : 1; d < D; ++d) acc += (u128)xf[d * SP + ks_slot(w)] * (uint64_t)kp[d * dstride + e * TL];
            x[e] = MONT ? redc((uint64_t)(acc >> 64), (uint64_t)acc, mi.M, qni)
                        : reduce128(mi, (uint64_t)(acc >> 64), (uint64_t)acc);
        }
    } else {
        const uint64_t q = mi.M;
        for (int e = 0; e < 16; ++e) {
            const int w = r16_wB<H>(e, lane);
            uint64_t s = 0;
            for (int d = 0; d < D; ++d) {
                const uint64_t p = mulmod(mi, reduce64(mi, xf[d * SP + ks_slot(w)]), (uint64_t)kp[d * dstride + e * TL]);
                s = s >= q - p ? s - (q - p) : s + p;
            }
            x[e] = MONT ? redc(0, s, q, qni) : s;
        }
    }
}

// K2 on the radix-16 register layouts (see k_ntt_pass_r16).  A workgroup
// takes the D digit columns of G limbs of one instance and one value `hi` of
// the high index bits (rows hi*W .. hi*W + W - 1):
//   forward  sub-problem s = d + D*gi on thread s + NF*lane (sub-problems
//            adjacent: a load covers adjacent columns of a row): load in
//            layout A, stages, LDS exchange, stages in layout B (reduced to
//            [0, q) unless the inner product takes lazy values), to LDS;
//   inverse  sub-problem s = c + 2*gi: the key inner product for the values
//            of layout B and the stages of layout B on thread lane + TL*s
//            (lanes adjacent: a key load covers TL adjacent words, whole
//            cache lines), LDS exchange, the stages of layout A and the
//            store on thread s + NI*lane (adjacent columns again).
// The prepared key holds window index wB(e, lane) of row block hi at
// hi*W + e*TL + lane.
// Tile `blk` of `nblk`, served by threads tid = 0.. of a workgroup with an
// LDS block lds of [max(NF, NI)][SP] slots.  A thread whose tid is beyond the
// tile's threads stays idle but runs the same barriers.
template <int H>
__device__ __forceinline__ void ks_tile(const InnerArgs& A, int64_t blk, int64_t nblk, int tid, uint64_t* lds) {
    constexpr int W = 1 << H;
    constexpr int TL = W / 16;
    constexpr int SP = ks_lds_row<H>();
    const int D = A.D, G = A.G;
    const int NF = G * D, NI = 2 * G;

    // instances vary fastest, then limb groups: the workgroups of one XCD share
    // their twiddles (same limbs and rows) and the partial cache lines at the
    // edges of neighbouring column groups of the same rows, in one L2
    const uint32_t tile = (uint32_t)xcd_tile(blk, nblk);
    const uint32_t tb = fdiv(tile, A.fd_B);
    const int64_t b = tile - tb * A.fd_B.d;
    const uint32_t thi = fdiv(tb, A.fd_ntg);
    const int64_t tg = tb - thi * A.fd_ntg.d;
    const int64_t hi = thi;
    const int64_t t0 = tg * G;
    const int64_t ncols1 = (int64_t)D * A.k * A.B, ncols2 = 2 * A.k * A.B;
    const int64_t row0 = hi * W;

    // ---- forward stages over the low bits (global stages hhi .. logm-1)
    {
        const bool act = tid < NF * TL;
        const int s = act ? tid % NF : 0, lane = act ? tid / NF : 0;
        int64_t t = t0 + s / D;
        const bool valid = act && t < A.k;
        if (t >= A.k) t = A.k - 1;
        const LimbInfo& L = A.limbs[t];
        const int mode = limb_mode(L);
        const ulonglong2* twrow = (const ulonglong2*)(A.tw_f + 2 * t * A.m);
        uint64_t x[16] = {};
        if (valid) {
            const uint64_t* src = A.src + (row0 + lane) * ncols1 + (s % D) + (int64_t)D * (t + A.k * b);
#pragma unroll
            for (int e = 0; e < 16; ++e) x[e] = src[(int64_t)(e * TL) * ncols1];
            r16_stages_any<H, false, 0>(x, lane, twrow, hi, A.logm, A.hhi, L, mode);
        }
        uint64_t* xs = lds + s * SP;
        if (act) {
#pragma unroll
            for (int e = 0; e < 16; ++e) xs[ks_slot(r16_wA<H>(e, lane))] = x[e];
        }
        __syncthreads();
        if (valid) {
#pragma unroll
            for (int e = 0; e < 16; ++e) x[e] = xs[ks_slot(r16_wB<H>(e, lane))];
            r16_stages_any<H, false, 1>(x, lane, twrow, hi, A.logm, A.hhi, L, mode);
            if (mode == MODE_L64) {
                const uint64_t q8 = L.q << 3, q4 = L.q << 2;
                if (ks_lazy_dot(L, D)) {
#pragma unroll
                    for (int e = 0; e < 16; ++e) {
                        const uint64_t v = x[e] >= q8 ? x[e] - q8 : x[e];
                        x[e] = v >= q4 ? v - q4 : v;
                    }
                } else {
#pragma unroll
                    for (int e = 0; e < 16; ++e) x[e] = lazy_reduce(x[e], L.q);
                }
            }
            // each thread rewrites exactly the slots it has just read
#pragma unroll
            for (int e = 0; e < 16; ++e) xs[ks_slot(r16_wB<H>(e, lane))] = x[e];
        }
        __syncthreads();
    }

    // ---- key inner product and sum over digits.  The chain computes
    //      prod_d = (T_DP(D_d) * T_DP(key_d)) % q      (modmul_ttt)
    //      sum    = (sum + prod_d) % q, d = 0..D-1     (axis_modsum)
    //      INTT input = residue of sum                  (intt load)
    // and the residue of sum is (sum_d D_d * key_d) mod q whatever the signs and
    // sizes of the key words, so the prepared key holds their residues and the
    // sum is accumulated exactly in 128 bits.
    const bool act = tid < NI * TL;
    uint64_t x[16] = {};
    {
        const int s = act ? tid / TL : 0, lane = act ? tid % TL : 0;
        const int c = s & 1, gi = s >> 1;
        int64_t t = t0 + gi;
        const bool valid = act && t < A.k;
        if (t >= A.k) t = A.k - 1;
        const LimbInfo& L = A.limbs[t];
        const int mode = limb_mode(L);
        if (valid) {
            const ModInfo mi = limb_modinfo(L);
            const bool single = ks_lazy_dot(L, D) || (uint64_t)D <= ~0ull / L.q;
            const uint64_t qni = ks_mont(L) ? L.qni : 0;
            const int64_t kb = A.key_b == 1 ? 0 : b;
            const int64_t* kp = A.key + ((kb * 2 + c) * A.key_k + t) * A.m + row0 + lane;
            const int64_t dstride = A.key_b * 2 * A.key_k * A.m;
            const uint64_t* xf = lds + (int64_t)(D * gi) * SP;
            if (qni) {
                switch (D) {
                    case 1: ks_dot<H, 1, true>(x, xf, kp, dstride, lane, mi, D, single, qni); break;
                    case 2: ks_dot<H, 2, true>(x, xf, kp, dstride, lane, mi, D, single, qni); break;
                    case 3: ks_dot<H, 3, true>(x, xf, kp, dstride, lane, mi, D, single, qni); break;
                    case 4: ks_dot<H, 4, true>(x, xf, kp, dstride, lane, mi, D, single, qni); break;
                    default: ks_dot<H, 0, true>(x, xf, kp, dstride, lane, mi, D, single, qni); break;
                }
            } else {
                switch (D) {
                    case 1: ks_dot<H, 1, false>(x, xf, kp, dstride, lane, mi, D, single, 0); break;
                    case 2: ks_dot<H, 2, false>(x, xf, kp, dstride, lane, mi, D, single, 0); break;
                    case 3: ks_dot<H, 3, false>(x, xf, kp, dstride, lane, mi, D, single, 0); break;
                    case 4: ks_dot<H, 4, false>(x, xf, kp, dstride, lane, mi, D, single, 0); break;
                    default: ks_dot<H, 0, false>(x, xf, kp, dstride, lane, mi, D, single, 0); break;
                }
            }
            // ---- inverse stages over the low bits (global stages 0 .. hlo-1), layout B
            const ulonglong2* twrow = (const ulonglong2*)(A.tw_i + 2 * t * A.m);
            r16_stages_any<H, true, 1>(x, lane, twrow, hi, A.logm, 0, L, mode);
        }
        __syncthreads();  // every read of the forward values is done
        if (act) {
            uint64_t* xs = lds + s * SP;
#pragma unroll
            for (int e = 0; e < 16; ++e) xs[ks_slot(r16_wB<H>(e, lane))] = x[e];
        }
        __syncthreads();
    }

    // ---- stages of layout A and the store, sub-problems on adjacent threads
    const int s = act ? tid % NI : 0, lane = act ? tid / NI : 0;
    const int c = s & 1, gi = s >> 1;
    const int64_t t = t0 + gi;
    if (!act || t >= A.k) return;
    const LimbInfo& L = A.limbs[t];
    const int mode = limb_mode(L);
    const uint64_t* xs = lds + s * SP;
#pragma unroll
    for (int e = 0; e < 16; ++e) x[e] = xs[ks_slot(r16_wA<H>(e, lane))];
    const ulonglong2* twrow = (const ulonglong2*)(A.tw_i + 2 * t * A.m);
    r16_stages_any<H, true, 0>(x, lane, twrow, hi, A.logm, 0, L, mode);

    // ---- store (residues stay lazy for the next radix-16 pass)
    uint64_t* dst = A.dst + (row0 + lane) * ncols2 + c + 2 * (t + A.k * b);
#pragma unroll
    for (int e = 0; e < 16; ++e) dst[(int64_t)(e * TL) * ncols2] = x[e];
}

template <int H>
__global__ __launch_bounds__(512) void k_ks_inner(InnerArgs A) {
    extern __shared__ uint64_t lds[];  // [max(NF, NI)][SP]
    ks_tile<H>(A, blockIdx.x, gridDim.x, threadIdx.x, lds);
}

// K2 of several key-switches of any inner window in one launch: workgroup b
// serves key-switch list[i] with b' = b - starts[i] and runs its tiles
// b' * per .. b' * per + per - 1 as `per` groups of `tpt` threads (members of
// smaller windows pack more tiles into a workgroup); threads beyond them, and
// tiles beyond the member's, stay idle but run every barrier.
__global__ __launch_bounds__(512, 3) void k_ks_inner_group(const InnerArgs* list, const int64_t* starts, int count) {
    extern __shared__ uint64_t lds[];
    const int i = instance_of(starts, count, blockIdx.x);
    const InnerArgs& A = list[i];
    const int64_t ntiles = A.n_hi * A.n_tg * A.B;
    const int sub = threadIdx.x / A.tpt;
    const int64_t tile = ((int64_t)blockIdx.x - starts[i]) * A.per + sub;
    const bool active = sub < A.per && tile < ntiles;
    const int tid = active ? (int)threadIdx.x % A.tpt : (1 << 30);
    const int rows = A.D * A.G > 2 * A.G ? A.D * A.G : 2 * A.G;
    const int64_t t0 = active ? tile : 0;
    const int H = A.H;
    uint64_t* base = lds + (int64_t)(active ? sub : 0) * rows * ks_lds_row_of(H);
    if (H == 8) ks_tile<8>(A, t0, ntiles, tid, base);
    else if (H == 7) ks_tile<7>(A, t0, ntiles, tid, base);
    else ks_tile<6>(A, t0, ntiles, tid, base);
}

// key[((d*B + b)*2 + c)*k + t][hi*W + e*TL + lane] = evk[d, b, c, rowmap[j], t] mod q_t
// (the residue in [0, q_t)) with j = hi*W + r16_wB(e, lane), W = 2^H the inner
// window (H = 6..8)
struct KeyArgs {
    const int64_t* evk;
    int64_t s[5];  // strides of the [D, B, 2, m, k] view
    int64_t* key;
    const int32_t* rowmap;  // null: identity
    const LimbInfo* limbs;  // key words are stored as residues mod q_t (Montgomery form where ks_mont)
    unsigned long long* negatives;  // count of negative key words
    int64_t D, B, m, k;
    int32_t H;
};

__global__ void k_ks_prepare_key(KeyArgs K) {
    const int64_t n = K.D * K.B * 2 * K.k * K.m;
    unsigned long long neg = 0;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t j = e % K.m, r = e / K.m;
        int64_t t = r % K.k;
        r /= K.k;
        int64_t c = r % 2;
        r /= 2;
        int64_t b = r % K.B, d = r / K.B;
        // position j inside its window block -> window index r16_wB(e, lane)
        const int nb = K.H - 4, tl = 1 << nb;
        const int64_t wmask = (int64_t(1) << K.H) - 1;
        const int pos = (int)(j & wmask), reg = pos / tl, lane = pos % tl;
        const int w = ((reg >> nb) << (2 * nb)) | (lane << nb) | (reg & ((1 << nb) - 1));
        const int64_t jw = (j & ~wmask) | w;
        int64_t u = K.rowmap ? (int64_t)K.rowmap[jw] : jw;
        const int64_t v = K.evk[d * K.s[0] + b * K.s[1] + c * K.s[2] + u * K.s[3] + t * K.s[4]];
        neg += v < 0;
        const LimbInfo& L = K.limbs[t];
        const ModInfo mi = limb_modinfo(L);
        const uint64_t res = residue(mi, v);
        // Montgomery form for the limbs whose inner product ends with a
        // Montgomery reduction: res * (2^64 mod q) mod q
        K.key[e] = (int64_t)(ks_mont(L) ? mulmod(mi, res, reduce128(mi, 1, 0)) : res);
    }
    if (neg) atomicAdd(K.negatives, neg);
}

// Galois maps: counts per (limb, row), then the inverse map galmap[t][perms[t, n]] = n.
__global__ void k_ks_gal_count(PlanSrc perms, int64_t k, int64_t m, int32_t* counts, unsigned long long* flags) {
    const int64_t n = k * m;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t t = e / m, u = e - t * m;
        int64_t v = ld_elem(perms.ptr, t * perms.s0 + u * perms.s1, perms.elem32);
        if (v < 0 || v >= m) {
            atomicAdd(&flags[0], 1ull);
            continue;
        }
        atomicAdd(&counts[t * m + v], 1);
    }
}

__global__ void k_ks_gal_map(PlanSrc perms, int64_t k, int64_t m, const int32_t* counts, int32_t* galmap,
                             unsigned long long* flags) {
    const int64_t n = k * m;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        if (counts[e] != 1) atomicAdd(&flags[1], 1ull);
        int64_t t = e / m, u = e - t * m;
        int64_t v = ld_elem(perms.ptr, t * perms.s0 + u * perms.s1, perms.elem32);
        if (v >= 0 && v < m) galmap[t * m + v] = (int32_t)u;
    }
}

// counts the entries of galmap [k][m] that differ from limb 0's row
__global__ void k_ks_gal_same(const int32_t* galmap, int64_t k, int64_t m, unsigned long long* flags) {
    for (int64_t e = m + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < k * m; e += (int64_t)gridDim.x * blockDim.x)
        if (galmap[e] != galmap[e % m]) atomicAdd(&flags[3], 1ull);
}

// forward and inverse plans must agree on moduli and permutation; flags[4]
// counts limbs where a digit (at most dmax) plus q can exceed 2^63 - 1, where
// the chain's forward NTT of the digits follows the reference's wrapping sums
__global__ void k_ks_compare(const LimbInfo* a, const LimbInfo* b, int64_t k, const int32_t* ma, const int32_t* mb,
                             int64_t m, uint64_t dmax, unsigned long long* flags) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < k + m; e += (int64_t)gridDim.x * blockDim.x) {
        if (e < k) {
            if (a[e].q != b[e].q) atomicAdd(&flags[2], 1ull);
            if (dmax > (~0ull >> 1) - a[e].q) atomicAdd(&flags[4], 1ull);
        } else if (ma && mb && ma[e - k] != mb[e - k]) {
            atomicAdd(&flags[2], 1ull);
        }
    }
}

}  // namespace

static void keyswitch_tune(hx_keyswitch_plan* plan, hipStream_t st);

hx_keyswitch_plan* keyswitch_plan_create(const hx_ntt_plan* fwd, const hx_ntt_plan* inv, const hx_view* perms,
                                         uint64_t dnum, uint64_t base_bits, hipStream_t caller) {
    if (!fwd || !inv) raise(HX_ERR_INVALID_ARGUMENT, "plan pointers must not be null.");
    if (fwd->inverse || !inv->inverse) raise(HX_ERR_INVALID_ARGUMENT, "need a forward and an inverse NTT plan");
    if (fwd->dtype != HX_INT64 || inv->dtype != HX_INT64)
        raise(HX_ERR_INVALID_ARGUMENT, "the fused key-switch works on int64 tensors");
    if (fwd->m != inv->m || fwd->k != inv->k) raise(HX_ERR_INVALID_ARGUMENT, "plans differ in m or k");
    if (fwd->device != inv->device) raise(HX_ERR_INVALID_ARGUMENT, "plans belong to different devices");
    int dev = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (dev != fwd->device) raise(HX_ERR_INVALID_ARGUMENT, "plans belong to another device");
    const int64_t m = fwd->m, k = fwd->k;
    const int logm = fwd->logm;
    if (logm < 12 || logm > 16) raise(HX_ERR_INVALID_ARGUMENT, "the fused key-switch needs 2^12 <= m <= 2^16");
    const bool fmap = fwd->perm_kind == PERM_MAP, imap = inv->perm_kind == PERM_MAP;
    if (!((fwd->perm_kind == PERM_IDENTITY || fmap) && fwd->perm_kind == inv->perm_kind))
        raise(HX_ERR_INVALID_ARGUMENT, "forward and inverse plans must share one permutation");
    if (dnum < 1 || dnum > 64) raise(HX_ERR_INVALID_ARGUMENT, "dnum must be in [1, 64]");
    if (base_bits >= 64) raise(HX_ERR_INVALID_ARGUMENT, "base_bits must be < 64");
    check_view(perms, "perms");
    if (perms->dtype != HX_INT32 && perms->dtype != HX_INT64) raise(HX_ERR_INVALID_ARGUMENT, "perms must be integer");
    if (perms->ndim != 2 || perms->dims[0] != k || perms->dims[1] != m)
        raise(HX_ERR_INVALID_ARGUMENT, "Perms must have shape [l, m] where l and m match a.shape at elementwise and perm axes.");
    if (fwd->ref_all || inv->ref_all)
        raise(HX_ERR_INVALID_ARGUMENT,
              "the fused key-switch needs NTT tables in [0, q) and moduli below 2^62 (use the six-call chain)");

    // preparation synchronises; inside a graph capture it runs on the side stream
    SyncScope sc(caller);
    hipStream_t st = sc.stream();
    hx_keyswitch_plan* plan = new hx_keyswitch_plan;
    std::memset(plan, 0, sizeof *plan);
    plan->device = dev;
    plan->fwd = fwd;
    plan->inv = inv;
    plan->m = m;
    plan->k = k;
    plan->logm = logm;
    plan->hlo = logm / 2;
    plan->hhi = logm - plan->hlo;
    plan->D = (int)dnum;
    plan->w = (int)base_bits;
    plan->G = 1;
    plan->tile = 32;
    const size_t map_bytes = ((size_t)k * m * 4 + 255) / 256 * 256;
    check_hip(hipMallocAsync(&plan->mem, 2 * map_bytes + 256, st), "hipMallocAsync(keyswitch plan)");
    plan->galmap = (int32_t*)plan->mem;
    int32_t* counts = (int32_t*)((char*)plan->mem + map_bytes);
    unsigned long long* flags = (unsigned long long*)((char*)plan->mem + 2 * map_bytes);
    plan->wss = new Workspaces;
    plan->wss->device = dev;
    try {
        fill_zero(counts, map_bytes + 256, st);
        PlanSrc ps = plan_src(perms);
        unsigned g = grid_for(k * m, 256, 8192);
        const uint64_t dmax = base_bits ? (~0ull >> (64 - base_bits)) : 0;
        hipLaunchKernelGGL(k_ks_gal_count, dim3(g), dim3(256), 0, st, ps, k, m, counts, flags);
        hipLaunchKernelGGL(k_ks_gal_map, dim3(g), dim3(256), 0, st, ps, k, m, counts, plan->galmap, flags);
        hipLaunchKernelGGL(k_ks_compare, dim3(grid_for(k + m, 256, 1024)), dim3(256), 0, st, fwd->limbs, inv->limbs, k,
                           fmap ? fwd->rowmap : nullptr, imap ? inv->rowmap : nullptr, m, dmax, flags);
        hipLaunchKernelGGL(k_ks_gal_same, dim3(g), dim3(256), 0, st, plan->galmap, k, m, flags);
        check_hip(hipGetLastError(), "keyswitch plan launch");
        unsigned long long h[5];
        check_hip(hipMemcpyAsync(h, flags, sizeof h, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
        sc.sync();
        if (h[0]) raise(HX_ERR_OUT_OF_RANGE, "Permutation index out of bounds.");
        if (h[1]) raise(HX_ERR_INVALID_ARGUMENT, "the fused key-switch needs every row of perms to be a permutation");
        if (h[2]) raise(HX_ERR_INVALID_ARGUMENT, "forward and inverse plans differ in moduli or permutation");
        if (h[4])
            raise(HX_ERR_INVALID_ARGUMENT,
                  "base_bits too large for these moduli: a digit plus q can exceed 2^63 - 1 (use the six-call chain)");
        plan->gal_shared = h[3] == 0;
        keyswitch_tune(plan, st);
    } catch (...) {
        (void)hipStreamSynchronize(st);
        std::vector<hipEvent_t> events;
        std::vector<void*> ptrs{plan->mem};
        plan->wss->release(events, ptrs);
        release_after(dev, events, ptrs);
        delete plan->wss;
        delete plan;
        throw;
    }
    return plan;
}

// freed after the work already queued with the plan (stream-ordered, no host wait)
void keyswitch_plan_destroy(hx_keyswitch_plan* plan) {
    if (!plan) return;
    std::vector<hipEvent_t> events;
    std::vector<void*> ptrs{plan->mem};
    {
        std::lock_guard<std::mutex> lk(plan->wss->mu);
        plan->wss->release(events, ptrs);
    }
    release_after(plan->device, events, ptrs);
    delete plan->wss;
    delete plan;
}

// Inverse of one Galois permutation perms [m] (or [1, m]) for keyswitch_exec's
// row_map: row_map[perms[n]] = n; perms must be a permutation of [0, m).
void keyswitch_row_map(const hx_keyswitch_plan* plan, const hx_view* perms, int32_t* row_map, hipStream_t caller) {
    if (!plan) raise(HX_ERR_INVALID_ARGUMENT, "plan pointer must not be null.");
    if (!row_map) raise(HX_ERR_INVALID_ARGUMENT, "row_map pointer must not be null.");
    check_view(perms, "perms");
    if (perms->dtype != HX_INT32 && perms->dtype != HX_INT64) raise(HX_ERR_INVALID_ARGUMENT, "perms must be integer");
    const int64_t m = plan->m;
    const bool ok1 = perms->ndim == 1 && perms->dims[0] == m;
    const bool ok2 = perms->ndim == 2 && perms->dims[0] == 1 && perms->dims[1] == m;
    if (!ok1 && !ok2) raise(HX_ERR_INVALID_ARGUMENT, "perms must have shape [m] or [1, m]");
    hx_view v = *perms;
    if (ok1) {  // as [1, m]
        v.ndim = 2;
        v.dims[1] = m;
        v.strides[1] = perms->strides[0];
        v.dims[0] = 1;
        v.strides[0] = 0;
    }
    const size_t cnt_bytes = ((size_t)m * 4 + 255) / 256 * 256;
    SyncScope sc(caller);  // validation synchronises; inside a capture it runs on the side stream
    hipStream_t st = sc.stream();
    void* mem = nullptr;
    check_hip(hipMallocAsync(&mem, cnt_bytes + 256, st), "hipMallocAsync(row map)");
    int32_t* counts = (int32_t*)mem;
    unsigned long long* flags = (unsigned long long*)((char*)mem + cnt_bytes);
    unsigned long long h[2] = {0, 0};
    try {
        fill_zero(mem, cnt_bytes + 256, st);
        PlanSrc ps = plan_src(&v);
        unsigned g = grid_for(m, 256, 8192);
        hipLaunchKernelGGL(k_ks_gal_count, dim3(g), dim3(256), 0, st, ps, (int64_t)1, m, counts, flags);
        hipLaunchKernelGGL(k_ks_gal_map, dim3(g), dim3(256), 0, st, ps, (int64_t)1, m, counts, row_map, flags);
        check_hip(hipGetLastError(), "row map launch");
        check_hip(hipMemcpyAsync(h, flags, sizeof h, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
        sc.sync();
    } catch (...) {
        (void)hipFreeAsync(mem, st);
        throw;
    }
    check_hip(hipFreeAsync(mem, st), "hipFreeAsync(row map)");
    if (h[0]) raise(HX_ERR_OUT_OF_RANGE, "Permutation index out of bounds.");
    if (h[1]) raise(HX_ERR_INVALID_ARGUMENT, "perms must be a permutation");
}

static void require_i64(const hx_view* v, const char* name, int ndim) {
    check_view(v, name);
    if (v->dtype != HX_INT64) raise(HX_ERR_INVALID_ARGUMENT, std::string(name) + " must be int64");
    if (v->ndim != ndim) raise(HX_ERR_INVALID_ARGUMENT, std::string(name) + " has the wrong number of dims");
}

size_t keyswitch_key_bytes(const hx_keyswitch_plan* plan, int64_t batch) {
    if (!plan) raise(HX_ERR_INVALID_ARGUMENT, "plan pointer must not be null.");
    if (batch < 1) raise(HX_ERR_INVALID_ARGUMENT, "batch must be positive");
    return (size_t)plan->D * (size_t)batch * 2 * (size_t)plan->k * (size_t)plan->m * 8;
}

// The key is checked as it is laid out (one synchronisation): the chain's
// modmul keeps the sign of a negative key word (truncating %), after which its
// INTT follows the reference's signed network, which the fused kernels do not
// reproduce; such keys are rejected and belong to the six-call chain.
void keyswitch_prepare_key(const hx_keyswitch_plan* plan, const hx_view* evk, void* key, hipStream_t caller) {
    if (!plan) raise(HX_ERR_INVALID_ARGUMENT, "plan pointer must not be null.");
    require_i64(evk, "evk", 5);
    if (!key) raise(HX_ERR_INVALID_ARGUMENT, "key pointer must not be null.");
    if (evk->dims[0] != plan->D || evk->dims[2] != 2 || evk->dims[3] != plan->m || evk->dims[4] != plan->k)
        raise(HX_ERR_INVALID_ARGUMENT, "evk must have shape [dnum, B, 2, m, k]");
    SyncScope sc(caller);
    hipStream_t st = sc.stream();
    unsigned long long* neg = nullptr;
    check_hip(hipMallocAsync((void**)&neg, 256, st), "hipMallocAsync");
    fill_zero(neg, 256, st);
    KeyArgs K;
    K.negatives = neg;
    K.evk = (const int64_t*)evk->data;
    for (int i = 0; i < 5; ++i) K.s[i] = evk->strides[i];
    K.key = (int64_t*)key;
    K.rowmap = plan->fwd->perm_kind == PERM_MAP ? plan->fwd->rowmap : nullptr;
    K.limbs = plan->fwd->limbs;
    K.D = plan->D;
    K.B = evk->dims[1];
    K.m = plan->m;
    K.k = plan->k;
    K.H = plan->hlo;
    hipLaunchKernelGGL(k_ks_prepare_key, dim3(grid_for(K.D * K.B * 2 * K.k * K.m, 256, 8192)), dim3(256), 0, st, K);
    unsigned long long h = 0;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&h, neg, sizeof h, hipMemcpyDeviceToHost, st);
    (void)hipFreeAsync(neg, st);
    check_hip(e, "prepare key launch");
    sc.sync();
    if (h)
        raise(HX_ERR_INVALID_ARGUMENT,
              "the fused key-switch needs key words >= 0: " + std::to_string(h) +
                  " negative word(s) (use the six-call chain)");
}

// The three launches of one key-switch, fully described.
struct KsArgs {
    PassArgs P1;  // K1
    InnerArgs A;  // K2
    PassArgs P3;  // K3
    int64_t blocks1, blocks2, blocks3;
    int H1, H2, threads2;
    size_t lds2;
};

// K2 launch shape for limb group G
static void inner_shape(int H, int D, int G, int* threads, size_t* lds) {
    const int TL = (1 << H) / 16;
    const int rows = std::max(G * D, 2 * G);
    *threads = (rows * TL + 63) / 64 * 64;
    *lds = (size_t)rows * ks_lds_row_of(H) * 8;
}

static bool inner_fits(int H, int D, int G) {
    int threads;
    size_t lds;
    inner_shape(H, D, G, &threads, &lds);
    return threads <= 512 && lds <= 64 * 1024;
}

// Workspace of a batch of B: the K1 -> K2 intermediate (D k B columns) and
// the K2 -> K3 one (2 k B columns), m rows each.
size_t keyswitch_ws_bytes(const hx_keyswitch_plan* plan, int64_t B) {
    return (size_t)plan->m * (size_t)((int64_t)plan->D * plan->k * B + 2 * plan->k * B) * 8;
}

// Validates one key-switch and fills the launch arguments, with the
// intermediates in ws (keyswitch_ws_bytes of the batch).  Returns false for an
// empty batch.
bool keyswitch_build(const hx_keyswitch_plan* plan, const hx_view* x, const void* key, size_t key_bytes,
                     int64_t key_limbs, int64_t key_count, const int32_t* row_map, const hx_view* out, void* ws,
                     KsArgs& K, int tile = 0) {
    if (plan && !tile) tile = plan->tile;
    if (!plan) raise(HX_ERR_INVALID_ARGUMENT, "plan pointer must not be null.");
    require_i64(x, "x", 3);
    require_i64(out, "out", 4);
    if (!key) raise(HX_ERR_INVALID_ARGUMENT, "key pointer must not be null.");
    const int64_t B = x->dims[0], m = plan->m, k = plan->k;
    const int D = plan->D;
    if (key_limbs < 0) key_limbs = k;
    if (key_count < 0) key_count = B;
    if (key_limbs < k) raise(HX_ERR_INVALID_ARGUMENT, "the key has fewer limbs than the plan");
    if (key_count != 1 && key_count != B) raise(HX_ERR_INVALID_ARGUMENT, "key_count must be 1 or the batch size");
    if (B > 0 && key_bytes < (size_t)D * (size_t)key_count * 2 * (size_t)key_limbs * (size_t)m * 8)
        raise(HX_ERR_INVALID_ARGUMENT, "key buffer smaller than a key for this batch");
    if (x->dims[1] != m || x->dims[2] != k) raise(HX_ERR_INVALID_ARGUMENT, "x must have shape [B, m, k]");
    if (out->dims[0] != B || out->dims[1] != 2 || out->dims[2] != m || out->dims[3] != k)
        raise(HX_ERR_INVALID_ARGUMENT, "out must have shape [B, 2, m, k]");
    if (views_overlap(x, out)) raise(HX_ERR_INVALID_ARGUMENT, "out must not overlap x");
    int dev = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (dev != plan->device) raise(HX_ERR_INVALID_ARGUMENT, "plan belongs to another device");
    if (B == 0) return false;
    if (!r16_window(plan->hhi) || !r16_window(plan->hlo))
        raise(HX_ERR_RUNTIME, "the fused key-switch needs the radix-16 pass kernels");
    const int64_t ncols1 = (int64_t)D * k * B, ncols2 = 2 * k * B;
    if (!ws) raise(HX_ERR_RUNTIME, "internal: key-switch workspace missing");
    uint64_t* tmp1 = (uint64_t*)ws;
    uint64_t* tmp2 = tmp1 + (size_t)m * ncols1;
    const hx_ntt_plan* F = plan->fwd;
    const hx_ntt_plan* I = plan->inv;

    // K1: digits of x, forward pass over the high bits
    PassArgs& P1 = K.P1;
    std::memset(&P1, 0, sizeof P1);
    P1.src.ptr = x->data;
    P1.src.row_stride = x->strides[1];
    P1.src.cstride[0] = 0;              // digit
    P1.src.cstride[1] = x->strides[2];  // limb
    P1.src.cstride[2] = x->strides[0];  // instance
    P1.src.kind = 3;
    P1.dst = tmp_addr(tmp1, ncols1);
    P1.ncols = ncols1;
    P1.cdim[0] = D;
    P1.cdim[1] = k;
    P1.cdim[2] = B;
    P1.limb_axis = 1;
    P1.gd_axis = 0;
    P1.gd_w = plan->w;
    P1.gd_D = D;
    P1.gd_status = &device_status()->gdecomp_overflow;
    P1.inverse = 0;
    P1.logm = plan->logm;
    P1.H = plan->hhi;
    P1.b0 = plan->hlo;
    P1.stage0 = 0;
    P1.bot_count = int64_t(1) << plan->hlo;
    P1.reduce_in = 1;
    P1.lazy_out = 1;
    P1.limbs = F->limbs;
    P1.tw = F->tw;
    P1.m = m;
    configure_tile(P1, true, 2, tile);
    K.blocks1 = pass_blocks(P1);
    K.H1 = plan->hhi;

    // K2: forward low bits, key inner product, inverse low bits
    InnerArgs& A = K.A;
    std::memset(&A, 0, sizeof A);
    A.src = tmp1;
    A.dst = tmp2;
    A.key = (const int64_t*)key;
    A.key_k = key_limbs;
    A.key_b = key_count;
    A.limbs = F->limbs;
    A.tw_f = F->tw;
    A.tw_i = I->tw;
    A.m = m;
    A.k = k;
    A.B = B;
    A.logm = plan->logm;
    A.hhi = plan->hhi;
    A.D = D;
    const int G = (int)std::min<int64_t>(plan->G, k);
    A.G = G;
    A.n_tg = (k + G - 1) / G;
    A.n_hi = int64_t(1) << plan->hhi;
    A.H = plan->hlo;
    K.H2 = plan->hlo;
    inner_shape(K.H2, D, G, &K.threads2, &K.lds2);
    if (K.threads2 > 512 || K.lds2 > 64 * 1024)
        raise(HX_ERR_INVALID_ARGUMENT, "dnum too large for the fused key-switch tile");
    K.blocks2 = A.n_hi * A.n_tg * B;
    if (K.blocks2 > 0x7fffffff) raise(HX_ERR_INVALID_ARGUMENT, "key-switch batch too large for one launch");
    A.fd_B = make_fastdiv((uint32_t)B);
    A.fd_ntg = make_fastdiv((uint32_t)A.n_tg);

    // K3: inverse pass over the high bits, m^-1, Galois scatter into out
    PassArgs& P3 = K.P3;
    std::memset(&P3, 0, sizeof P3);
    P3.src = tmp_addr(tmp2, ncols2);
    P3.dst.ptr = out->data;
    P3.dst.row_stride = out->strides[2];
    P3.dst.cstride[0] = out->strides[1];  // output polynomial c
    P3.dst.cstride[1] = out->strides[3];  // limb
    P3.dst.cstride[2] = out->strides[0];  // instance
    P3.dst.kind = 0;
    P3.ncols = ncols2;
    P3.cdim[0] = 2;
    P3.cdim[1] = k;
    P3.cdim[2] = B;
    P3.limb_axis = 1;
    P3.inverse = 1;
    P3.logm = plan->logm;
    P3.H = plan->hhi;
    P3.b0 = plan->hlo;
    P3.stage0 = plan->hlo;
    P3.bot_count = int64_t(1) << plan->hlo;
    P3.scale_out = 1;
    if (row_map) P3.dst_map = row_map;  // one Galois permutation for every limb
    else if (plan->gal_shared) P3.dst_map = plan->galmap;  // limb 0's row serves every limb
    else P3.dst_map_limb = plan->galmap;
    P3.limbs = I->limbs;
    P3.tw = I->tw;
    P3.m = m;
    configure_tile(P3, true, 2, tile);
    K.blocks3 = pass_blocks(P3);
    return true;
}

static void launch_inner(const KsArgs& K, hipStream_t st) {
    const dim3 grid((unsigned)K.blocks2), block((unsigned)K.threads2);
    if (K.H2 == 8) hipLaunchKernelGGL(k_ks_inner<8>, grid, block, K.lds2, st, K.A);
    else if (K.H2 == 7) hipLaunchKernelGGL(k_ks_inner<7>, grid, block, K.lds2, st, K.A);
    else hipLaunchKernelGGL(k_ks_inner<6>, grid, block, K.lds2, st, K.A);
    check_hip(hipGetLastError(), "keyswitch inner launch");
}

// marks (may be null): events recorded before K1, after K1, after K2 and after K3
static void keyswitch_launch(KsArgs& K, hipStream_t st, void* const* marks = nullptr) {
    auto mark = [&](int i) {
        if (marks && marks[i]) check_hip(hipEventRecord((hipEvent_t)marks[i], st), "hipEventRecord");
    };
    mark(0);
    launch_pass(K.P1, st);
    mark(1);
    launch_inner(K, st);
    mark(2);
    launch_pass(K.P3, st);
    mark(3);
}

void keyswitch_exec(const hx_keyswitch_plan* plan, const hx_view* x, const void* key, size_t key_bytes,
                    int64_t key_limbs, int64_t key_count, const int32_t* row_map, const hx_view* out, hipStream_t st,
                    void* const* marks = nullptr, const RmacArgs* rm = nullptr) {
    if (!plan) raise(HX_ERR_INVALID_ARGUMENT, "plan pointer must not be null.");
    require_i64(x, "x", 3);
    if (x->dims[0] == 0) return;
    std::lock_guard<std::mutex> lk(plan->wss->mu);
    const Workspaces::Ws w = plan->wss->acquire(st, 0, keyswitch_ws_bytes(plan, x->dims[0]));
    KsArgs K;
    if (keyswitch_build(plan, x, key, key_bytes, key_limbs, key_count, row_map, out, w.data, K)) {
        if (rm) K.P3.rm = *rm;
        keyswitch_launch(K, st, marks);
    }
    plan->wss->note(st);
}

// The key-switch of ct's second polynomial with the rotate-multiply-accumulate
// of the serving loop in its last kernel (RmacArgs): acc is updated in place
// and the key-switch output is never written.  acc and ct are [B, 2, m, k]
// with equal strides, diags [B, M, M, k] contiguous per instance (M a power of
// two); the moduli are the plan's.
void keyswitch_exec_rmac(const hx_keyswitch_plan* plan, const hx_view* x, const void* key, size_t key_bytes,
                         int64_t key_limbs, int64_t key_count, const int32_t* row_map, const hx_view* acc,
                         const hx_view* ct, const hx_view* diags, int64_t diag, hipStream_t st) {
    if (!plan) raise(HX_ERR_INVALID_ARGUMENT, "plan pointer must not be null.");
    require_i64(acc, "acc", 4);
    require_i64(ct, "ct", 4);
    require_i64(diags, "diags", 4);
    for (int d = 0; d < 4; ++d) {
        if (ct->dims[d] != acc->dims[d] || (ct->dims[d] > 1 && ct->strides[d] != acc->strides[d]))
            raise(HX_ERR_INVALID_ARGUMENT, "ct must have the shape and strides of acc");
    }
    const int64_t B = acc->dims[0], M = diags->dims[1], k = plan->k;
    if (diags->dims[0] != B || diags->dims[2] != M || diags->dims[3] != k)
        raise(HX_ERR_INVALID_ARGUMENT, "diags must have shape [B, M, M, k]");
    if (M <= 0 || (M & (M - 1)) != 0) raise(HX_ERR_INVALID_ARGUMENT, "the matrix size M must be a power of two");
    if ((k > 1 && diags->strides[3] != 1) || (M > 1 && diags->strides[2] != k) || (M > 1 && diags->strides[1] != M * k))
        raise(HX_ERR_INVALID_ARGUMENT, "diags must be contiguous within an instance");
    if (diag < 0) raise(HX_ERR_INVALID_ARGUMENT, "diag must be >= 0");
    if (views_overlap(acc, ct) || views_overlap(acc, diags) || views_overlap(acc, x))
        raise(HX_ERR_INVALID_ARGUMENT, "acc must not overlap the other operands");
    RmacArgs rm;
    rm.on = 1;
    rm.ct = (const int64_t*)ct->data;
    rm.diags = (const int64_t*)diags->data;
    rm.diags_b = diags->strides[0];
    rm.M = M;
    rm.diag = diag % M;
    rm.L = k;
    keyswitch_exec(plan, x, key, key_bytes, key_limbs, key_count, row_map, acc, st, nullptr, &rm);
}

// ------------------------------------------------------------ autotuning
// Per ring degree m and dnum: the split of log m into the K1/K3 window (high
// bits) and the K2 window (low bits); per (m, k, dnum): the K2 limb group G
// and the K1/K3 tile (32 or 16 columns per workgroup).
// Every choice computes the same result; they differ in speed only.  The split
// is chosen once per (m, dnum) in a process, so keys prepared by one plan serve
// every plan of that ring degree.  Choices are kept in a text file keyed by
// (GPU architecture, m, k, dnum) — HEAL_AMD_TUNE_CACHE, else
// $HOME/.cache/heal_amd/keyswitch_tune.txt — so that later processes skip the
// timing; a file that cannot be read or written is ignored.  One lock covers
// lookup, timing and recording, so concurrent plan creation never pairs a
// split with a group tuned for another split.
static std::mutex g_tune_mu;
static std::map<std::pair<int64_t, int>, int> g_tuned_hlo;
static std::map<std::tuple<int64_t, int64_t, int>, int> g_tuned_group;
static std::map<std::tuple<int64_t, int64_t, int>, int> g_tuned_tile;
static bool g_tune_file_read = false;

static std::string tune_cache_path() {
    if (const char* e = getenv("HEAL_AMD_TUNE_CACHE")) return e;
    const char* home = getenv("HOME");
    return home ? std::string(home) + "/.cache/heal_amd/keyswitch_tune.txt" : std::string();
}

static std::string gpu_arch(int dev) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
    std::string a = prop.gcnArchName;
    const size_t c = a.find(':');
    return c == std::string::npos ? a : a.substr(0, c);
}

// Lines "arch m k dnum low_bits limb_group tile" (a line without the tile
// means 32); entries of other architectures, malformed lines and entries a
// process has already decided are skipped.  Caller holds g_tune_mu.
static void tune_file_read(const std::string& arch) {
    if (g_tune_file_read) return;
    g_tune_file_read = true;
    const std::string path = tune_cache_path();
    if (path.empty()) return;
    FILE* f = fopen(path.c_str(), "r");
    if (!f) return;
    char line[256];
    while (fgets(line, sizeof line, f)) {
        char a[64];
        long long m, k;
        int D, hlo, G, tile = 32;
        const int n = sscanf(line, "%63s %lld %lld %d %d %d %d", a, &m, &k, &D, &hlo, &G, &tile);
        if (n < 6 || arch != a || hlo < 6 || hlo > 8 || G < 1 || G > 8 || (tile != 16 && tile != 32)) continue;
        g_tuned_hlo.emplace(std::make_pair((int64_t)m, D), hlo);
        // a group and a tile are kept only with the split they were timed for
        if (g_tuned_hlo[{(int64_t)m, D}] == hlo) {
            g_tuned_group.emplace(std::make_tuple((int64_t)m, (int64_t)k, D), G);
            g_tuned_tile.emplace(std::make_tuple((int64_t)m, (int64_t)k, D), tile);
        }
    }
    fclose(f);
}

static void tune_file_append(const std::string& arch, int64_t m, int64_t k, int D, int hlo, int G, int tile) {
    const std::string path = tune_cache_path();
    if (path.empty()) return;
    // best effort: create the directories on the path
    for (size_t i = path.find('/', 1); i != std::string::npos; i = path.find('/', i + 1))
        (void)mkdir(path.substr(0, i).c_str(), 0755);
    FILE* f = fopen(path.c_str(), "a");
    if (!f) return;
    fprintf(f, "%s %lld %lld %d %d %d %d\n", arch.c_str(), (long long)m, (long long)k, D, hlo, G, tile);
    fclose(f);
}

// Fastest of five timed key-switches after one warm-up (the minimum is the
// least noisy estimate of a kernel sequence's cost on a shared machine).
static float time_keyswitch(hx_keyswitch_plan* plan, const hx_view* x, void* key, size_t key_bytes, const hx_view* out,
                            void* ws, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    KsArgs K;
    keyswitch_build(plan, x, key, key_bytes, -1, -1, nullptr, out, ws, K);
    keyswitch_launch(K, st);  // warm-up
    float best = 1e30f;
    for (int r = 0; r < 5; ++r) {
        check_hip(hipEventRecord(e0, st), "hipEventRecord");
        keyswitch_launch(K, st);
        check_hip(hipEventRecord(e1, st), "hipEventRecord");
        check_hip(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0;
        check_hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        best = std::min(best, ms);
    }
    return best;
}

static void keyswitch_tune(hx_keyswitch_plan* plan, hipStream_t st) {
    const int64_t m = plan->m, k = plan->k;
    const int D = plan->D, logm = plan->logm;
    std::vector<int> hlos;
    for (int h = 6; h <= 8; ++h)
        if (logm - h >= 6 && logm - h <= 8 && inner_fits(h, D, 1)) hlos.push_back(h);
    std::vector<int> groups;
    for (int G = 1; G <= 8; G *= 2)
        if (G <= std::max<int64_t>(k, 1)) groups.push_back(G);
    const int default_hlo = logm / 2;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const std::string arch = gpu_arch(plan->device);
    tune_file_read(arch);
    int hlo = -1, G = -1, tile = 32;
    {
        auto a = g_tuned_hlo.find({m, D});
        if (a != g_tuned_hlo.end() && std::find(hlos.begin(), hlos.end(), a->second) != hlos.end()) hlo = a->second;
        auto b = g_tuned_group.find({m, k, D});
        if (hlo >= 0 && b != g_tuned_group.end() && inner_fits(hlo, D, b->second)) G = b->second;
        auto c = g_tuned_tile.find({m, k, D});
        if (c != g_tuned_tile.end()) tile = c->second;
    }
    if (hlo >= 0 && G >= 0) {
        plan->hlo = hlo;
        plan->hhi = logm - hlo;
        plan->G = G;
        plan->tile = tile;
        return;
    }
    // scratch key-switch: B instances of zeros (the timing does not depend on
    // values), enough instances that every candidate fills the GPU: at
    // m * k = 2^16 * 24 that is 8 instances (about 0.9 GB of scratch)
    const int64_t B = std::max<int64_t>(1, std::min<int64_t>(8, (int64_t(1) << 24) / (m * k)));
    const size_t xb = (size_t)B * m * k * 8, ob = 2 * xb, kb = (size_t)D * B * 2 * k * m * 8;
    const size_t wb = keyswitch_ws_bytes(plan, B);
    void* mem = nullptr;
    check_hip(hipMallocAsync(&mem, xb + ob + kb + wb, st), "hipMallocAsync(tuning)");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    try {
        fill_zero(mem, xb + ob + kb, st);
        check_hip(hipEventCreate(&e0), "hipEventCreate");
        check_hip(hipEventCreate(&e1), "hipEventCreate");
        hx_view xv, ov;
        std::memset(&xv, 0, sizeof xv);
        std::memset(&ov, 0, sizeof ov);
        xv.data = mem;
        xv.dtype = HX_INT64;
        xv.ndim = 3;
        xv.dims[0] = B, xv.dims[1] = m, xv.dims[2] = k;
        xv.strides[0] = m * k, xv.strides[1] = k, xv.strides[2] = 1;
        ov.data = (char*)mem + xb;
        ov.dtype = HX_INT64;
        ov.ndim = 4;
        ov.dims[0] = B, ov.dims[1] = 2, ov.dims[2] = m, ov.dims[3] = k;
        ov.strides[0] = 2 * m * k, ov.strides[1] = m * k, ov.strides[2] = k, ov.strides[3] = 1;
        void* key = (char*)mem + xb + ob;
        void* ws = (char*)mem + xb + ob + kb;
        plan->tile = 32;
        if (hlo < 0) {
            float best = 1e30f;
            for (int h : hlos) {
                plan->hlo = h;
                plan->hhi = logm - h;
                plan->G = std::max(1, std::min(8, 12 / D));
                while (plan->G > 1 && !inner_fits(h, D, plan->G)) plan->G /= 2;
                const float t = time_keyswitch(plan, &xv, key, kb, &ov, ws, st, e0, e1);
                if (t < best || (t == best && h == default_hlo)) best = t, hlo = h;
            }
        }
        plan->hlo = hlo;
        plan->hhi = logm - hlo;
        float best = 1e30f;
        for (int g : groups) {
            if (!inner_fits(hlo, D, g)) continue;
            plan->G = g;
            const float t = time_keyswitch(plan, &xv, key, kb, &ov, ws, st, e0, e1);
            if (t < best) best = t, G = g;
        }
        // K1/K3 workgroups of 16 sub-problems (twice as many workgroups, each
        // half the size) against the 32 timed above: taken only when clearly
        // faster, since the two time within noise of each other where both
        // fill the GPU and 32 is faster in the accumulate-fused last kernel
        if (G >= 0) {
            plan->G = G;
            plan->tile = 16;
            if (time_keyswitch(plan, &xv, key, kb, &ov, ws, st, e0, e1) < 0.97f * best) tile = 16;
            plan->tile = 32;
        }
    } catch (...) {
        (void)hipStreamSynchronize(st);
        (void)hipFreeAsync(mem, st);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    check_hip(hipFreeAsync(mem, st), "hipFreeAsync(tuning)");
    check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (G < 0 || !inner_fits(hlo, D, G)) raise(HX_ERR_INVALID_ARGUMENT, "dnum too large for the fused key-switch tile");
    plan->G = G;
    plan->tile = tile;
    g_tuned_hlo[{m, D}] = hlo;
    g_tuned_group[{m, k, D}] = G;
    g_tuned_tile[{m, k, D}] = tile;
    tune_file_append(arch, m, k, D, hlo, G, tile);
}

// ------------------------------------------------------------ grouped execution
// Several key-switches of any (m, k) with fixed buffers: per phase, one launch
// per window size over device arrays of launch arguments.
struct GroupLaunch {
    int phase;  // 1, 2, 3
    int H;
    int count;
    int threads;
    size_t lds;
    int64_t blocks;
    size_t args_off, starts_off;  // byte offsets in the group's device memory
};

}  // namespace hx

struct hx_keyswitch_group {
    int device;
    void* mem;                       // launch arguments and tile offsets
    std::vector<void*> workspaces;   // one per member (stream-ordered allocations)
    std::vector<hx::GroupLaunch> launches;
    hx::Workspaces uses;             // the streams that ran the group (events only)
};

namespace hx {

hx_keyswitch_group* keyswitch_group_create(const hx_keyswitch_plan* const* plans, const hx_view* xs,
                                           const void* const* keys, const size_t* key_bytes, const hx_view* outs,
                                           int count, hipStream_t caller) {
    if (count < 1) raise(HX_ERR_INVALID_ARGUMENT, "a group needs at least one key-switch");
    if (!plans || !xs || !keys || !key_bytes || !outs) raise(HX_ERR_INVALID_ARGUMENT, "array pointers must not be null.");
    // every member gets a workspace of its own (owned by the group), so members
    // may share a plan and the plans stay free for other work
    SyncScope sc(caller);
    hipStream_t st = sc.stream();
    auto* g = new hx_keyswitch_group;
    g->mem = nullptr;
    std::vector<KsArgs> ks;
    auto drop = [&]() {
        (void)hipStreamSynchronize(st);
        std::vector<hipEvent_t> none;
        std::vector<void*> ptrs = g->workspaces;
        ptrs.push_back(g->mem);
        release_after(g->device, none, ptrs);
        delete g;
    };
    try {
        check_hip(hipGetDevice(&g->device), "hipGetDevice");
        g->uses.device = g->device;
        for (int i = 0; i < count; ++i) {
            if (!plans[i]) raise(HX_ERR_INVALID_ARGUMENT, "plan pointer must not be null.");
            require_i64(&xs[i], "x", 3);
            void* ws = nullptr;
            if (xs[i].dims[0] > 0) {
                check_hip(hipMallocAsync(&ws, keyswitch_ws_bytes(plans[i], xs[i].dims[0]), st), "hipMallocAsync(group)");
                g->workspaces.push_back(ws);
            }
            KsArgs K;
            // the group kernels serve 32-sub-problem tiles whatever the plan's tuned tile
            const bool any =
                keyswitch_build(plans[i], &xs[i], keys[i], key_bytes[i], -1, -1, nullptr, &outs[i], ws, K, kR16Tile);
            if (any) ks.push_back(K);
        }
    } catch (...) {
        drop();
        throw;
    }
    // host image: per launch [args...][starts...]; one launch per phase for
    // every member whatever its window sizes
    std::vector<char> image;
    auto align = [&](size_t a) { image.resize((image.size() + a - 1) / a * a); };
    for (int phase = 1; phase <= 3 && !ks.empty(); ++phase) {
        GroupLaunch L;
        L.phase = phase;
        L.H = 0;
        L.count = (int)ks.size();
        L.threads = phase == 2 ? 0 : 512;
        L.lds = 0;
        align(256);
        L.args_off = image.size();
        const size_t asz = phase == 2 ? sizeof(InnerArgs) : sizeof(PassArgs);
        image.resize(image.size() + asz * ks.size());
        std::vector<int64_t> starts(1, 0);
        // K2: every member packs as many tiles into a workgroup as the widest
        // member's threads allow (and 64 KB of LDS)
        int k2_threads = 64;
        if (phase == 2)
            for (const KsArgs& K : ks) k2_threads = std::max(k2_threads, K.threads2);
        if (phase == 2) L.threads = k2_threads;
        for (size_t i = 0; i < ks.size(); ++i) {
            KsArgs K = ks[i];
            int64_t wgs;
            if (phase == 2) {
                const int TL = (1 << K.H2) / 16;
                const int rows = std::max((int)K.A.D * K.A.G, 2 * K.A.G);
                const size_t tile_lds = (size_t)rows * ks_lds_row_of(K.H2) * 8;
                K.A.tpt = rows * TL;
                K.A.per = std::max(1, k2_threads / K.A.tpt);
                while (K.A.per > 1 && (size_t)K.A.per * tile_lds > 64 * 1024) --K.A.per;
                wgs = (K.blocks2 + K.A.per - 1) / K.A.per;
                L.lds = std::max(L.lds, (size_t)K.A.per * tile_lds);
            } else {
                const int H = phase == 1 ? K.P1.H : K.P3.H;
                const int per = 1 << (8 - H);
                wgs = ((phase == 1 ? K.blocks1 : K.blocks3) + per - 1) / per;
                L.lds = std::max(L.lds, (size_t)per * kR16Tile * ((1 << H) + 1) * 8);
            }
            const void* src = phase == 1 ? (const void*)&K.P1 : (phase == 2 ? (const void*)&K.A : (const void*)&K.P3);
            std::memcpy(image.data() + L.args_off + i * asz, src, asz);
            starts.push_back(starts.back() + wgs);
        }
        L.blocks = starts.back();
        align(256);
        L.starts_off = image.size();
        image.resize(image.size() + starts.size() * sizeof(int64_t));
        std::memcpy(image.data() + L.starts_off, starts.data(), starts.size() * sizeof(int64_t));
        g->launches.push_back(L);
    }
    try {
        check_hip(hipMallocAsync(&g->mem, std::max<size_t>(image.size(), 256), st), "hipMallocAsync(group)");
        check_hip(hipMemcpyAsync(g->mem, image.data(), image.size(), hipMemcpyHostToDevice, st), "hipMemcpyAsync(group)");
        sc.sync();
    } catch (...) {
        drop();
        throw;
    }
    return g;
}

void keyswitch_group_exec(const hx_keyswitch_group* g, hipStream_t st) {
    if (!g) raise(HX_ERR_INVALID_ARGUMENT, "group pointer must not be null.");
    int dev = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (dev != g->device) raise(HX_ERR_INVALID_ARGUMENT, "group belongs to another device");
    std::lock_guard<std::mutex> lk(const_cast<hx_keyswitch_group*>(g)->uses.mu);
    for (const GroupLaunch& L : g->launches) {
        const char* base = (const char*)g->mem;
        const int64_t* starts = (const int64_t*)(base + L.starts_off);
        if (L.blocks > 0x7fffffff) raise(HX_ERR_INVALID_ARGUMENT, "too many tiles for one launch");
        const dim3 grid((unsigned)L.blocks), block((unsigned)L.threads);
        if (L.phase == 2) {
            hipLaunchKernelGGL(k_ks_inner_group, grid, block, L.lds, st, (const InnerArgs*)(base + L.args_off), starts,
                               L.count);
        } else if (L.phase == 3) {
            hipLaunchKernelGGL(k_ntt_pass_r16_group<true>, grid, block, L.lds, st,
                               (const PassArgs*)(base + L.args_off), starts, L.count);
        } else {
            hipLaunchKernelGGL(k_ntt_pass_r16_group<false>, grid, block, L.lds, st,
                               (const PassArgs*)(base + L.args_off), starts, L.count);
        }
        check_hip(hipGetLastError(), "grouped keyswitch launch");
    }
    const_cast<hx_keyswitch_group*>(g)->uses.note(st);
}

// freed after the work already queued with the group (stream-ordered, no host wait)
void keyswitch_group_destroy(hx_keyswitch_group* g) {
    if (!g) return;
    std::vector<hipEvent_t> events;
    std::vector<void*> ptrs = g->workspaces;
    ptrs.push_back(g->mem);
    {
        std::lock_guard<std::mutex> lk(g->uses.mu);
        g->uses.release(events, ptrs);
    }
    release_after(g->device, events, ptrs);
    delete g;
}

}  // namespace hx

using namespace hx;

extern "C" {

int hx_keyswitch_plan_create(hx_keyswitch_plan** out, const hx_ntt_plan* fwd, const hx_ntt_plan* inv,
                             const hx_view* perms, uint64_t dnum, uint64_t base_bits, hx_stream stream) {
    HX_API_BEGIN
    if (!out) raise(HX_ERR_INVALID_ARGUMENT, "out pointer must not be null.");
    *out = nullptr;
    *out = keyswitch_plan_create(fwd, inv, perms, dnum, base_bits, as_stream(stream));
    HX_API_END
}

int hx_keyswitch_key_bytes(const hx_keyswitch_plan* plan, int64_t batch, size_t* bytes) {
    HX_API_BEGIN
    if (!bytes) raise(HX_ERR_INVALID_ARGUMENT, "bytes pointer must not be null.");
    *bytes = keyswitch_key_bytes(plan, batch);
    HX_API_END
}

int hx_keyswitch_plan_config(const hx_keyswitch_plan* plan, int32_t* low_bits, int32_t* limb_group, int32_t* tile) {
    HX_API_BEGIN
    if (!plan) raise(HX_ERR_INVALID_ARGUMENT, "plan pointer must not be null.");
    if (low_bits) *low_bits = plan->hlo;
    if (limb_group) *limb_group = plan->G;
    if (tile) *tile = plan->tile;
    HX_API_END
}

int hx_keyswitch_prepare_key(const hx_keyswitch_plan* plan, const hx_view* evk, void* key, hx_stream stream) {
    HX_API_BEGIN
    keyswitch_prepare_key(plan, evk, key, as_stream(stream));
    HX_API_END
}

int hx_keyswitch_exec(const hx_keyswitch_plan* plan, const hx_view* x, const void* key, size_t key_bytes,
                      const hx_view* out, hx_stream stream) {
    HX_API_BEGIN
    keyswitch_exec(plan, x, key, key_bytes, -1, -1, nullptr, out, as_stream(stream));
    HX_API_END
}

int hx_keyswitch_exec_marked(const hx_keyswitch_plan* plan, const hx_view* x, const void* key, size_t key_bytes,
                             const hx_view* out, void* const* events, hx_stream stream) {
    HX_API_BEGIN
    if (!events) raise(HX_ERR_INVALID_ARGUMENT, "events pointer must not be null.");
    keyswitch_exec(plan, x, key, key_bytes, -1, -1, nullptr, out, as_stream(stream), events);
    HX_API_END
}

int hx_keyswitch_exec_ex(const hx_keyswitch_plan* plan, const hx_view* x, const void* key, size_t key_bytes,
                         int64_t key_limbs, int64_t key_count, const int32_t* row_map, const hx_view* out,
                         hx_stream stream) {
    HX_API_BEGIN
    if (key_limbs < 1 || key_count < 1) raise(HX_ERR_INVALID_ARGUMENT, "key_limbs and key_count must be positive");
    keyswitch_exec(plan, x, key, key_bytes, key_limbs, key_count, row_map, out, as_stream(stream));
    HX_API_END
}

int hx_keyswitch_exec_rmac(const hx_keyswitch_plan* plan, const hx_view* x, const void* key, size_t key_bytes,
                           int64_t key_limbs, int64_t key_count, const int32_t* row_map, const hx_view* acc,
                           const hx_view* ct, const hx_view* diags, int64_t diag, hx_stream stream) {
    HX_API_BEGIN
    if (key_limbs < 1 || key_count < 1) raise(HX_ERR_INVALID_ARGUMENT, "key_limbs and key_count must be positive");
    keyswitch_exec_rmac(plan, x, key, key_bytes, key_limbs, key_count, row_map, acc, ct, diags, diag,
                        as_stream(stream));
    HX_API_END
}

int hx_keyswitch_group_create(hx_keyswitch_group** out, const hx_keyswitch_plan* const* plans, const hx_view* xs,
                              const void* const* keys, const size_t* key_bytes, const hx_view* outs, int count,
                              hx_stream stream) {
    HX_API_BEGIN
    if (!out) raise(HX_ERR_INVALID_ARGUMENT, "out pointer must not be null.");
    *out = nullptr;
    *out = keyswitch_group_create(plans, xs, keys, key_bytes, outs, count, as_stream(stream));
    HX_API_END
}

int hx_keyswitch_group_exec(const hx_keyswitch_group* group, hx_stream stream) {
    HX_API_BEGIN
    keyswitch_group_exec(group, as_stream(stream));
    HX_API_END
}

int hx_keyswitch_group_destroy(hx_keyswitch_group* group) {
    HX_API_BEGIN
    keyswitch_group_destroy(group);
    HX_API_END
}

int hx_keyswitch_row_map(const hx_keyswitch_plan* plan, const hx_view* perms, int32_t* row_map, hx_stream stream) {
    HX_API_BEGIN
    keyswitch_row_map(plan, perms, row_map, as_stream(stream));
    HX_API_END
}

int hx_keyswitch_plan_destroy(hx_keyswitch_plan* plan) {
    HX_API_BEGIN
    keyswitch_plan_destroy(plan);
    HX_API_END
}

}  // extern "C"
