// hx_arith.hpp — exact 64-bit modular arithmetic for gfx950 (and the host).
//
// The reference computes every modular product with a double-width integer
// and `%` (example_impl/typing.h:8-20: int32 -> int64, int64 -> __int128).
// CDNA4 has no 128-bit divide, so this file reduces with precomputed
// reciprocals instead:
//
//  * General moduli (modops, on-load reduction): division by an invariant
//    64-bit divisor after Möller & Granlund, "Improved division by invariant
//    integers", IEEE Trans. Computers 60(2), 2011 (algorithm "div_2by1"): the
//    divisor is normalised (top bit set) and its reciprocal
//    v = floor((2^128-1)/d) - 2^64 is computed once per modulus.
//  * Fixed multiplicands (NTT twiddles, m^-1): Shoup's method, V. Shoup, NTL
//    "MulModPrecon": w' = floor(w * 2^64 / q); for x < 2^64,
//    r = x*w - floor(x*w'/2^64)*q lies in [0, 2q), so one conditional
//    subtraction gives the exact residue.  Needs q < 2^63.
//
// All results are the exact residues, so they are bit-identical with the
// reference's 128-bit `%` wherever the reference itself is defined.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define HX_HD __host__ __device__ __forceinline__
#else
#define HX_HD inline
#endif

namespace hx {

typedef unsigned __int128 u128;

HX_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((u128)a * b) >> 64);
#endif
}

HX_HD int clz64(uint64_t x) {  // x != 0
#if defined(__HIP_DEVICE_COMPILE__)
    return __clzll((long long)x);
#else
    return __builtin_clzll(x);
#endif
}

// Reduction data for one modulus M (1 <= M < 2^64): d = M << s has its top bit
// set and v is the Möller–Granlund reciprocal of d.
struct ModInfo {
    uint64_t M;
    uint64_t d;
    uint64_t v;
    uint32_t s;
    uint32_t pad;
};

// floor(((2^64-1-d) * 2^64 + 2^64-1) / d) for normalised d, by restoring
// long division (64 steps).  Used once per modulus.
HX_HD uint64_t mg_reciprocal(uint64_t d) {
    uint64_t r = ~d;
    uint64_t q = 0;
    for (int i = 63; i >= 0; --i) {
        uint64_t carry = r >> 63;
        r = (r << 1) | 1u;  // the low word of the numerator is all ones
        q <<= 1;
        if (carry || r >= d) {
            r -= d;
            q |= 1u;
        }
    }
    return q;
}

HX_HD ModInfo make_modinfo(uint64_t M) {
    ModInfo mi;
    mi.M = M;
    mi.s = (uint32_t)clz64(M);
    mi.d = M << mi.s;
    mi.v = mg_reciprocal(mi.d);
    mi.pad = 0;
    return mi;
}

// Remainder of (u1:u0) / d for normalised d with reciprocal v; requires u1 < d.
// Also returns the quotient word through *qout when non-null.
HX_HD uint64_t mg_divrem(uint64_t u1, uint64_t u0, uint64_t d, uint64_t v, uint64_t* qout) {
    u128 q = (u128)v * u1;
    q += ((u128)(u1 + 1) << 64) | u0;
    uint64_t q1 = (uint64_t)(q >> 64);
    uint64_t q0 = (uint64_t)q;
    uint64_t r = u0 - q1 * d;
    if (r > q0) {
        q1 -= 1;
        r += d;
    }
    if (r >= d) {
        q1 += 1;
        r -= d;
    }
    if (qout) *qout = q1;
    return r;
}

// (hi:lo) mod M; requires (hi:lo) < M * 2^64 (always true for hi < M).
HX_HD uint64_t reduce128(const ModInfo& mi, uint64_t hi, uint64_t lo) {
    uint32_t s = mi.s;
    uint64_t u1 = s ? ((hi << s) | (lo >> (64 - s))) : hi;
    uint64_t u0 = lo << s;
    return mg_divrem(u1, u0, mi.d, mi.v, nullptr) >> s;
}

HX_HD uint64_t reduce64(const ModInfo& mi, uint64_t x) {
    if (x < mi.M) return x;
    return reduce128(mi, 0, x);
}

// a * b mod M for a, b < M.
HX_HD uint64_t mulmod(const ModInfo& mi, uint64_t a, uint64_t b) {
    u128 p = (u128)a * b;
    return reduce128(mi, (uint64_t)(p >> 64), (uint64_t)p);
}

// floor((hi:lo) / M); requires hi < M.
HX_HD uint64_t divq128(const ModInfo& mi, uint64_t hi, uint64_t lo) {
    uint32_t s = mi.s;
    uint64_t u1 = s ? ((hi << s) | (lo >> (64 - s))) : hi;
    uint64_t u0 = lo << s;
    uint64_t q1;
    mg_divrem(u1, u0, mi.d, mi.v, &q1);
    return q1;
}

// Shoup companion floor(w * 2^64 / q) for w < q.
HX_HD uint64_t shoup_companion(const ModInfo& mi, uint64_t w) {
    uint32_t s = mi.s;
    uint64_t q1;
    // (w << s) : 0 divided by (q << s) equals floor(w * 2^64 / q); w < q keeps the quotient one word.
    mg_divrem(w << s, 0, mi.d, mi.v, &q1);
    return q1;
}

// x * w mod q in [0, 2q) for any x < 2^64, w < q < 2^63, wp = shoup_companion(w).
HX_HD uint64_t shoup_lazy(uint64_t x, uint64_t w, uint64_t wp, uint64_t q) {
    uint64_t qh = mulhi64(x, wp);
    return x * w - qh * q;
}

// shoup_lazy with nq = 2^64 - q: x*w + qh*nq = x*w - qh*q (mod 2^64), written
// over 32-bit halves so that the two low products share one 64-bit
// multiply-add chain and the cross terms one high word.
HX_HD uint64_t shoup_lazy_nq(uint64_t x, uint64_t w, uint64_t wp, uint64_t nq) {
    const uint64_t qh = mulhi64(x, wp);
    const uint32_t xl = (uint32_t)x, xh = (uint32_t)(x >> 32);
    const uint32_t wl = (uint32_t)w, wh = (uint32_t)(w >> 32);
    const uint32_t ql = (uint32_t)qh, qhh = (uint32_t)(qh >> 32);
    const uint32_t nl = (uint32_t)nq, nh = (uint32_t)(nq >> 32);
    const uint64_t p = (uint64_t)xl * wl + (uint64_t)ql * nl;  // mod 2^64
    const uint32_t hi = (uint32_t)(p >> 32) + xl * wh + xh * wl + ql * nh + qhh * nl;
    return ((uint64_t)hi << 32) | (uint32_t)p;
}

HX_HD uint32_t mulhi32(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umulhi(a, b);
#else
    return (uint32_t)(((uint64_t)a * b) >> 32);
#endif
}

// shoup_lazy_nq with a three-product quotient.  With x = xh:xl and
// wp = ph:pl in 32-bit halves,
//     floor(x*wp / 2^64) = xh*ph + floor(xh*pl / 2^32) + floor(xl*ph / 2^32) + c,
// where c in {0, 1, 2} collects the low words of the two cross products and
// the high word of xl*pl.  Leaving c out gives a quotient at most 2 below
// Shoup's, so r = x*w - qh*q lies in [0, 4q) for any x < 2^64 and
// w < q < 2^62 (Shoup's own remainder lies in [0, 2q)).  It saves the
// fourth partial product and the carry chain of the exact high product.
HX_HD uint64_t shoup_lazy3_nq(uint64_t x, uint64_t w, uint64_t wp, uint64_t nq) {
    const uint32_t xl = (uint32_t)x, xh = (uint32_t)(x >> 32);
    const uint32_t pl = (uint32_t)wp, ph = (uint32_t)(wp >> 32);
    const uint64_t qh = (uint64_t)xh * ph + mulhi32(xh, pl) + mulhi32(xl, ph);
    const uint32_t wl = (uint32_t)w, wh = (uint32_t)(w >> 32);
    const uint32_t ql = (uint32_t)qh, qhh = (uint32_t)(qh >> 32);
    const uint32_t nl = (uint32_t)nq, nh = (uint32_t)(nq >> 32);
    const uint64_t p = (uint64_t)xl * wl + (uint64_t)ql * nl;  // mod 2^64
    const uint32_t hi = (uint32_t)(p >> 32) + xl * wh + xh * wl + ql * nh + qhh * nl;
    return ((uint64_t)hi << 32) | (uint32_t)p;
}

// Lazy radix-2 butterflies for 64-bit moduli q < 2^60 (D. Harvey, "Faster
// arithmetic for number-theoretic transforms", J. Symb. Comp. 60, 2014) on the
// three-product remainder t = shoup_lazy3_nq(x1, w, wp, -q) in [0, 4q):
//   forward (Cooley-Tukey)  x0 <- u + t, x1 <- u - t + 4q, u = x0 reduced by a
//       conditional subtraction of 8q when RED.  RED: x0 < 16q gives u < 8q and
//       outputs below 12q; without RED x0 < 12q gives outputs below 16q <= 2^64.
//       Alternating the two (a reducing stage first) keeps every value below 16q.
//   inverse (Gentleman-Sande)  x0 <- (x0 + x1) reduced below 4q, x1 <- (x0 - x1 +
//       4q) w, for x0, x1 < 4q.
template <bool RED>
HX_HD void lazy_fwd_bfly(uint64_t& x0, uint64_t& x1, uint64_t w, uint64_t wp, uint64_t q) {
    uint64_t u = x0;
    if (RED) {
        const uint64_t q8 = q << 3;
        u = x0 >= q8 ? x0 - q8 : x0;
    }
    const uint64_t t = shoup_lazy3_nq(x1, w, wp, (uint64_t)0 - q);
    const uint64_t q4 = q << 2;
    x0 = u + t;
    x1 = u - t + q4;
}

HX_HD void lazy_inv_bfly(uint64_t& x0, uint64_t& x1, uint64_t w, uint64_t wp, uint64_t q) {
    const uint64_t q4 = q << 2;
    const uint64_t a = x0 + x1;
    const uint64_t d = x0 - x1 + q4;
    x0 = a >= q4 ? a - q4 : a;
    x1 = shoup_lazy3_nq(d, w, wp, (uint64_t)0 - q);
}

// The residue in [0, q) of a lazy value x < 16q.
HX_HD uint64_t lazy_reduce(uint64_t x, uint64_t q) {
    const uint64_t q8 = q << 3, q4 = q << 2, q2 = q << 1;
    x = x >= q8 ? x - q8 : x;
    x = x >= q4 ? x - q4 : x;
    x = x >= q2 ? x - q2 : x;
    return x >= q ? x - q : x;
}

// Montgomery reduction with R = 2^64 (P. L. Montgomery, "Modular multiplication
// without trial division", Math. Comp. 44, 1985).  For odd q: qni = -q^-1 mod
// 2^64 by Newton's iteration x <- x (2 - q x), each step doubling the correct
// low bits (q x == 1 mod 8 to start).
HX_HD uint64_t neg_inv64(uint64_t q) {
    uint64_t x = q;  // q * q == 1 mod 8 for odd q
    for (int i = 0; i < 5; ++i) x *= 2 - q * x;
    return (uint64_t)0 - x;
}

// (hi:lo) * 2^-64 mod q, in [0, 2q), for odd q < 2^63 and (hi:lo) < q * 2^64.
// With m = lo * qni mod 2^64, lo + m q == 0 mod 2^64, so (hi:lo) + m q is
// hi + mulhi(m, q) + [lo != 0] times 2^64.
HX_HD uint64_t redc(uint64_t hi, uint64_t lo, uint64_t q, uint64_t qni) {
    const uint64_t m = lo * qni;
    return hi + mulhi64(m, q) + (lo != 0 ? 1u : 0u);
}

HX_HD uint64_t shoup_mul(uint64_t x, uint64_t w, uint64_t wp, uint64_t q) {
    uint64_t r = shoup_lazy(x, w, wp, q);
    return r >= q ? r - q : r;
}

// 32-bit Shoup for q < 2^31: wp = floor(w * 2^32 / q).
HX_HD uint32_t shoup_mul32(uint32_t x, uint32_t w, uint32_t wp, uint32_t q) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t qh = __umulhi(x, wp);
#else
    uint32_t qh = (uint32_t)(((uint64_t)x * wp) >> 32);
#endif
    uint32_t r = x * w - qh * q;
    return r >= q ? r - q : r;
}

HX_HD uint64_t addmod(uint64_t a, uint64_t b, uint64_t q) {  // a, b < q < 2^63
    uint64_t s = a + b;
    return s >= q ? s - q : s;
}

HX_HD uint64_t submod(uint64_t a, uint64_t b, uint64_t q) {  // a, b < q
    return a >= b ? a - b : a + q - b;
}

// ---------------------------------------------------------------------------
// C++ integer semantics of the reference's elementwise ops.
//
// The reference evaluates `static_cast<T>(T_DP(a) op T_DP(b)) % T_DP(p))`
// (example_impl/modop_impl.cpp:123-126): the double-width value is exact and
// `%` truncates toward zero, so the result is sign(x) * (|x| mod |p|).

HX_HD uint64_t uabs64(int64_t x) { return x < 0 ? (uint64_t)0 - (uint64_t)x : (uint64_t)x; }

HX_HD int64_t apply_sign(bool neg, uint64_t mag) { return neg ? (int64_t)((uint64_t)0 - mag) : (int64_t)mag; }

// (T_DP(a) * T_DP(b)) % T_DP(p) with |p| described by mi (mi.M = |p| > 0).
HX_HD int64_t cpp_mulmod(const ModInfo& mi, int64_t a, int64_t b) {
    bool neg = (a < 0) != (b < 0);
    uint64_t ma = reduce64(mi, uabs64(a));
    uint64_t mb = reduce64(mi, uabs64(b));
    uint64_t r = mulmod(mi, ma, mb);
    return apply_sign(neg && r != 0, r);
}

// (T_DP(a) + T_DP(b)) % T_DP(p).
HX_HD int64_t cpp_addmod(const ModInfo& mi, int64_t a, int64_t b) {
    bool na = a < 0, nb = b < 0;
    if (na == nb) {
        uint64_t ra = reduce64(mi, uabs64(a));
        uint64_t rb = reduce64(mi, uabs64(b));
        uint64_t r = ra + rb;  // < 2M, may wrap only if M > 2^63
        if (r < ra || r >= mi.M) r -= mi.M;
        return apply_sign(na && r != 0, r);
    }
    int64_t s = a + b;  // opposite signs: exact in 64 bits
    uint64_t r = reduce64(mi, uabs64(s));
    return apply_sign(s < 0 && r != 0, r);
}

// a % b (C++ truncation) for |b| described by mi.
HX_HD int64_t cpp_rem(const ModInfo& mi, int64_t a) {
    uint64_t r = reduce64(mi, uabs64(a));
    return apply_sign(a < 0 && r != 0, r);
}

}  // namespace hx
