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
    if (failures) {
        std::printf("FAILED %d of %ld\n", failures, checks);
        return 1;
    }
    std::printf("OK %ld\n", checks);
    return 0;
}
