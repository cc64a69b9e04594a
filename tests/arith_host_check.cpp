// Host-side check of heal_amd/csrc/hx_arith.hpp against the compiler's exact
// 128-bit `%` (the reference's T_DP arithmetic, example_impl/typing.h:8-20).
// Built and run by tests/test_arith_host.py; prints "OK <count>" on success.
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <random>
#include "../heal_amd/csrc/hx_arith.hpp"

using hx::u128;
typedef __int128 i128;

static int failures = 0;
static long checks = 0;

#define CHECK(cond, ...)                                  \
    do {                                                  \
        ++checks;                                         \
        if (!(cond)) {                                    \
            if (failures++ < 20) {                        \
                std::printf("FAIL line %d: ", __LINE__);  \
                std::printf(__VA_ARGS__);                 \
                std::printf("\n");                        \
            }                                             \
        }                                                 \
    } while (0)

static uint64_t pick_mod(std::mt19937_64& rng) {
    switch (rng() % 6) {
        case 0: return 1 + rng() % 1000;
        case 1: return (rng() >> (rng() % 64)) | 1;
        case 2: return (rng() >> 1) | (1ull << 62);              // 63-bit
        case 3: return (rng() >> 4) | 1;                          // 60-bit
        case 4: return (1ull << (rng() % 63));                     // powers of two
        default: return (rng() >> 33) + 1;                         // 31-bit
    }
}

static int64_t pick_val(std::mt19937_64& rng, uint64_t M) {
    switch (rng() % 5) {
        case 0: return (int64_t)(rng() % M);
        case 1: return (int64_t)rng();
        case 2: return -(int64_t)(rng() % M);
        case 3: return (rng() & 1) ? INT64_MIN : INT64_MAX;
        default: return (int64_t)(rng() >> (rng() % 64)) * ((rng() & 1) ? 1 : -1);
    }
}

int main() {
    std::mt19937_64 rng(12345);
    for (int it = 0; it < 200000; ++it) {
        uint64_t M = pick_mod(rng);
        if (M == 0) M = 1;
        hx::ModInfo mi = hx::make_modinfo(M);
        // reciprocal definition
        u128 num = ~(u128)0;
        u128 expect_v = num / mi.d - ((u128)1 << 64);
        CHECK((u128)mi.v == expect_v, "reciprocal M=%llu", (unsigned long long)M);
        // reduce128 on hi < M
        uint64_t hi = rng() % M, lo = rng();
        u128 x = ((u128)hi << 64) | lo;
        CHECK(hx::reduce128(mi, hi, lo) == (uint64_t)(x % M), "reduce128 M=%llu", (unsigned long long)M);
        uint64_t y = rng();
        CHECK(hx::reduce64(mi, y) == y % M, "reduce64");
        uint64_t a = rng() % M, b = rng() % M;
        CHECK(hx::mulmod(mi, a, b) == (uint64_t)(((u128)a * b) % M), "mulmod");
        if (M < (1ull << 63) && M > 1) {
            uint64_t w = rng() % M;
            uint64_t wp = hx::shoup_companion(mi, w);
            CHECK((u128)wp == (((u128)w << 64) / M), "shoup companion");
            uint64_t xx = rng();
            CHECK(hx::shoup_mul(xx, w, wp, M) == (uint64_t)(((u128)xx * w) % M), "shoup_mul");
            if (M < (1ull << 62)) {
                // three-product quotient: congruent to x*w, below 4M, for any x < 2^64
                const uint64_t xs[4] = {xx, ~0ull, (xx | 0xffffffffull), M < (1ull << 61) ? rng() % (8 * M) : rng()};
                for (uint64_t xv : xs) {
                    const uint64_t ws[2] = {w, M - 1};
                    for (uint64_t wv : ws) {
                        const uint64_t wq = hx::shoup_companion(mi, wv);
                        const uint64_t r = hx::shoup_lazy3_nq(xv, wv, wq, (uint64_t)0 - M);
                        CHECK(r < 4 * M && r % M == (uint64_t)(((u128)xv * wv) % M), "shoup_lazy3 x=%llu w=%llu M=%llu",
                              (unsigned long long)xv, (unsigned long long)wv, (unsigned long long)M);
                    }
                }
            }
        }
        if ((M & 1) && M > 1 && M < (1ull << 63)) {
            // Montgomery: redc(T) * 2^64 == T (mod M), redc(T) < 2M, for T < M * 2^64 (2M fits 64 bits)
            const uint64_t qni = hx::neg_inv64(M);
            CHECK(M * ((uint64_t)0 - qni) == 1, "neg_inv64 M=%llu", (unsigned long long)M);
            const uint64_t th[3] = {rng() % M, M - 1, 0};
            const uint64_t tl[3] = {rng(), ~0ull, rng() % 2};
            for (int i = 0; i < 3; ++i) {
                const u128 T = ((u128)th[i] << 64) | tl[i];
                const uint64_t r = hx::redc(th[i], tl[i], M, qni);
                CHECK(r < 2 * M, "redc bound M=%llu", (unsigned long long)M);
                CHECK((((u128)r << 64) % M) == T % M, "redc M=%llu hi=%llu lo=%llu", (unsigned long long)M,
                      (unsigned long long)th[i], (unsigned long long)tl[i]);
            }
        }
        if (M < (1ull << 31) && M > 1) {
            uint32_t q = (uint32_t)M, w = (uint32_t)(rng() % M), xx = (uint32_t)rng();
            uint32_t wp = (uint32_t)(((uint64_t)w << 32) / q);
            CHECK(hx::shoup_mul32(xx, w, wp, q) == (uint32_t)(((uint64_t)xx * w) % q), "shoup32");
        }
        // C++ semantics with signed operands: modulus p = +-M (|p| <= 2^63)
        if (M <= (1ull << 63)) {
            int64_t p = (M == (1ull << 63)) ? INT64_MIN : ((rng() & 1) ? -(int64_t)M : (int64_t)M);
            int64_t sa = pick_val(rng, M), sb = pick_val(rng, M);
            i128 em = ((i128)sa * (i128)sb) % (i128)p;
            i128 es = ((i128)sa + (i128)sb) % (i128)p;
            i128 er = (i128)sa % (i128)p;
            CHECK(hx::cpp_mulmod(mi, sa, sb) == (int64_t)em, "cpp_mulmod a=%lld b=%lld p=%lld", (long long)sa, (long long)sb, (long long)p);
            CHECK(hx::cpp_addmod(mi, sa, sb) == (int64_t)es, "cpp_addmod a=%lld b=%lld p=%lld", (long long)sa, (long long)sb, (long long)p);
            CHECK(hx::cpp_rem(mi, sa) == (int64_t)er, "cpp_rem");
        }
    }
    // 60-bit lazy butterflies (q < 2^60): bounds of each stage kind at the extremes
    // of their input ranges, congruence with exact arithmetic, and chains of
    // alternating stages (a reducing stage first, as every pass starts)
    for (int it = 0; it < 20000; ++it) {
        uint64_t q = (it % 4 == 0) ? ((1ull << 60) - 1 - 2 * (rng() % 4096)) : ((rng() >> 4) | 1);
        if (q < 3) q = 3;
        const hx::ModInfo mi = hx::make_modinfo(q);
        const uint64_t w = (it % 7 == 0) ? q - 1 : rng() % q;
        const uint64_t wp = hx::shoup_companion(mi, w);
        const uint64_t x1s[4] = {rng(), ~0ull, rng() % (16 * q), 0};
        auto mulw = [&](uint64_t x) { return (uint64_t)(((u128)(x % q) * w) % q); };
        // reducing stage: x0 < 16q -> outputs < 12q
        const uint64_t r0s[5] = {16 * q - 1, 8 * q, 8 * q - 1, 0, rng() % (16 * q)};
        for (uint64_t a : r0s)
            for (uint64_t b : x1s) {
                uint64_t x0 = a, x1 = b;
                hx::lazy_fwd_bfly<true>(x0, x1, w, wp, q);
                CHECK(x0 < 12 * q && x1 < 12 * q, "fwd RED bound q=%llu", (unsigned long long)q);
                CHECK(x0 % q == (a % q + mulw(b)) % q && x1 % q == (a % q + q - mulw(b)) % q, "fwd RED value");
            }
        // non-reducing stage: x0 < 12q -> outputs < 16q
        const uint64_t n0s[4] = {12 * q - 1, 0, rng() % (12 * q), 4 * q};
        for (uint64_t a : n0s)
            for (uint64_t b : x1s) {
                uint64_t x0 = a, x1 = b;
                hx::lazy_fwd_bfly<false>(x0, x1, w, wp, q);
                CHECK(x0 < 16 * q && x1 < 16 * q, "fwd bound q=%llu", (unsigned long long)q);
                CHECK(x0 % q == (a % q + mulw(b)) % q && x1 % q == (a % q + q - mulw(b)) % q, "fwd value");
            }
        // inverse: x0, x1 < 4q -> both < 4q
        const uint64_t i0s[3] = {4 * q - 1, 0, rng() % (4 * q)};
        for (uint64_t a : i0s)
            for (uint64_t b : i0s) {
                uint64_t x0 = a, x1 = b;
                hx::lazy_inv_bfly(x0, x1, w, wp, q);
                CHECK(x0 < 4 * q && x1 < 4 * q, "inv bound q=%llu", (unsigned long long)q);
                CHECK(x0 % q == (a + b) % q && x1 % q == mulw((a % q + q - b % q) % q), "inv value");
            }
        const uint64_t lz[4] = {16 * q - 1, 8 * q, rng() % (16 * q), q};
        for (uint64_t a : lz) CHECK(hx::lazy_reduce(a, q) == a % q, "lazy_reduce q=%llu", (unsigned long long)q);
        // a 16-point network of 4 stages twice over (8 stages, pairs i, i ^ 2^(s mod 4)), each
        // stage's twiddles random; residues tracked exactly alongside
        if (it % 8 == 0) {
            uint64_t x[16], e[16];
            for (int i = 0; i < 16; ++i) {
                x[i] = (i == 0) ? 16 * q - 1 : rng() % (16 * q);
                e[i] = x[i] % q;
            }
            for (int st = 0; st < 8; ++st) {
                const int h = 1 << (3 - st % 4);
                const uint64_t ws = rng() % q, wsp = hx::shoup_companion(mi, ws);
                for (int i = 0; i < 16; ++i) {
                    if (i & h) continue;
                    if (st % 2 == 0) hx::lazy_fwd_bfly<true>(x[i], x[i + h], ws, wsp, q);
                    else hx::lazy_fwd_bfly<false>(x[i], x[i + h], ws, wsp, q);
                    const uint64_t t = (uint64_t)(((u128)e[i + h] * ws) % q);
                    const uint64_t u = e[i];
                    e[i] = (u + t) % q;
                    e[i + h] = (u + q - t) % q;
                }
                for (int i = 0; i < 16; ++i)
                    CHECK(x[i] < (st % 2 == 0 ? 12 * q : 16 * q) && x[i] % q == e[i], "chain stage %d q=%llu", st,
                          (unsigned long long)q);
            }
        }
    }
    if (failures) {
        std::printf("FAILED %d of %ld\n", failures, checks);
        return 1;
    }
    std::printf("OK %ld\n", checks);
    return 0;
}
