// ref_harness.cpp — extern "C" entry points over the compiled reference
// (Lattica-ai/heal example_impl, built from its sources by oracle/Makefile into
// oracle/_ref/).  TEST INFRASTRUCTURE ONLY: used by tests/ to cross-check the
// C oracle and by bench.py's cpu_baseline leg; the product never loads it.
//
// Every function takes hx_view descriptors over host memory (the same
// signatures as heal_oracle.c), wraps them in torch tensors, runs the
// reference API (include/*.h of the reference) and copies the result back.
// Exceptions map to the status codes of include/heal_amd.h.
#include <chrono>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include <torch/torch.h>

#include "lattica_hw_api.h"
#include "../include/heal_amd.h"

using namespace lattica_hw_api;

namespace {

std::string g_err;

torch::Tensor wrap(const hx_view* v) {
    std::vector<int64_t> dims(v->dims, v->dims + v->ndim), strides(v->strides, v->strides + v->ndim);
    auto dt = v->dtype == HX_INT32 ? torch::kInt32 : (v->dtype == HX_INT64 ? torch::kInt64 : torch::kFloat64);
    return torch::from_blob(v->data, dims, strides, torch::TensorOptions().dtype(dt));
}

template <typename T>
std::shared_ptr<DeviceTensor<T>> up(const hx_view* v) {
    if (!v) return nullptr;
    return host_to_device<T>(wrap(v));
}

template <typename T>
void down(const std::shared_ptr<DeviceTensor<T>>& d, const hx_view* v) {
    torch::Tensor h = device_to_host<T>(d);
    wrap(v).copy_(h);
}

template <typename F>
int guard(F&& f) {
    try {
        f();
        return HX_OK;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return HX_ERR_INVALID_ARGUMENT;
    } catch (const std::out_of_range& e) {
        g_err = e.what();
        return HX_ERR_OUT_OF_RANGE;
    } catch (const std::bad_alloc& e) {
        g_err = e.what();
        return HX_ERR_BAD_ALLOC;
    } catch (const std::exception& e) {
        g_err = e.what();
        return HX_ERR_RUNTIME;
    }
}

template <typename T>
int modop_t(int op, const hx_view* a, const hx_view* b, int64_t bs, const hx_view* p, int64_t ps, const hx_view* r) {
    return guard([&] {
        auto A = up<T>(a), B = up<T>(b), P = up<T>(p), R = up<T>(r);
        if (op == 0) {
            if (b && p) modmul_ttt<T>(A, B, P, R);
            else if (b) modmul_ttc<T>(A, B, (T)ps, R);
            else if (p) modmul_tct<T>(A, (T)bs, P, R);
            else modmul_tcc<T>(A, (T)bs, (T)ps, R);
        } else {
            if (b && p) modsum_ttt<T>(A, B, P, R);
            else if (b) modsum_ttc<T>(A, B, (T)ps, R);
            else if (p) modsum_tct<T>(A, (T)bs, P, R);
            else modsum_tcc<T>(A, (T)bs, (T)ps, R);
        }
        down<T>(R, r);
    });
}

}  // namespace

extern "C" {

const char* ref_last_error(void) { return g_err.c_str(); }

int ref_modop(int op, const hx_view* a, const hx_view* b, int64_t bs, const hx_view* p, int64_t ps, const hx_view* r) {
    return r->dtype == HX_INT32 ? modop_t<int32_t>(op, a, b, bs, p, ps, r) : modop_t<int64_t>(op, a, b, bs, p, ps, r);
}

// variant: 0 = mod_tt, 1 = mod_tc, 2 = mod_ct; a null view passes a null tensor
int ref_mod(int variant, const hx_view* a, int64_t as, const hx_view* b, int64_t bs, const hx_view* r) {
    auto body = [&](auto tag) {
        using T = decltype(tag);
        return guard([&] {
            auto A = up<T>(a), B = up<T>(b), R = up<T>(r);
            if (variant == 0) mod_tt<T>(A, B, R);
            else if (variant == 1) mod_tc<T>(A, bs, R);
            else mod_ct<T>(as, B, R);
            down<T>(R, r);
        });
    };
    return r->dtype == HX_INT32 ? body(int32_t{}) : body(int64_t{});
}

int ref_axis_modsum(const hx_view* a, const hx_view* p, const hx_view* r, int64_t axis) {
    auto body = [&](auto tag) {
        using T = decltype(tag);
        return guard([&] {
            auto A = up<T>(a), P = up<T>(p), R = up<T>(r);
            axis_modsum<T>(A, P, R, axis);
            down<T>(R, r);
        });
    };
    return r->dtype == HX_INT32 ? body(int32_t{}) : body(int64_t{});
}

int ref_g_decomposition(const hx_view* a, const hx_view* r, uint64_t power, uint64_t base_bits) {
    auto body = [&](auto tag) {
        using T = decltype(tag);
        return guard([&] {
            auto A = up<T>(a), R = up<T>(r);
            g_decomposition<T>(A, R, power, base_bits);
            down<T>(R, r);
        });
    };
    return r->dtype == HX_INT32 ? body(int32_t{}) : body(int64_t{});
}

int ref_ntt(const hx_view* a, const hx_view* p, const hx_view* perm, const hx_view* tw, const hx_view* r) {
    auto body = [&](auto tag) {
        using T = decltype(tag);
        return guard([&] {
            auto A = up<T>(a), P = up<T>(p), PM = up<T>(perm), TW = up<T>(tw), R = up<T>(r);
            ntt<T>(A, P, PM, TW, nullptr, nullptr, R);
            down<T>(R, r);
        });
    };
    return r->dtype == HX_INT32 ? body(int32_t{}) : body(int64_t{});
}

int ref_intt(const hx_view* a, const hx_view* p, const hx_view* perm, const hx_view* itw, const hx_view* minv,
             const hx_view* r) {
    auto body = [&](auto tag) {
        using T = decltype(tag);
        return guard([&] {
            auto A = up<T>(a), P = up<T>(p), PM = up<T>(perm), TW = up<T>(itw), MI = up<T>(minv), R = up<T>(r);
            intt<T>(A, P, PM, TW, MI, nullptr, nullptr, R);
            down<T>(R, r);
        });
    };
    return r->dtype == HX_INT32 ? body(int32_t{}) : body(int64_t{});
}

int ref_permute(const hx_view* a, const hx_view* perms, const hx_view* r, int64_t ew, int64_t pa) {
    auto body = [&](auto tag) {
        using T = decltype(tag);
        return guard([&] {
            auto A = up<T>(a), PM = up<T>(perms), R = up<T>(r);
            permute<T>(A, PM, R, ew, pa);
            down<T>(R, r);
        });
    };
    return r->dtype == HX_INT32 ? body(int32_t{}) : body(int64_t{});
}

int ref_make_contiguous(const hx_view* src, const hx_view* dst) {
    auto body = [&](auto tag) {
        using T = decltype(tag);
        return guard([&] {
            auto S = up<T>(src);
            auto C = make_contiguous<T>(S);
            down<T>(C, dst);
        });
    };
    if (src->dtype == HX_INT32) return body(int32_t{});
    if (src->dtype == HX_INT64) return body(int64_t{});
    return body(double{});
}

// One hybrid key-switch through the reference API, the chain
// g_decomposition -> ntt -> modmul_ttt -> axis_modsum -> intt -> permute
// on int64 data (N coefficients, L limbs, dnum digits of base_bits):
//   x [N, L], evk [dnum, 2, N, L], p [L], perm [N], tw/itw [L, N], minv [L],
//   perms [L, N]; out [2, N, L].
// Only the six API calls are timed; *seconds receives their wall time.
int ref_keyswitch(const hx_view* x, const hx_view* evk, const hx_view* p, const hx_view* perm, const hx_view* tw,
                  const hx_view* itw, const hx_view* minv, const hx_view* perms, uint64_t dnum, uint64_t base_bits,
                  const hx_view* out, double* seconds) {
    return guard([&] {
        using T = int64_t;
        const int64_t N = x->dims[0], L = x->dims[1], D = (int64_t)dnum;
        auto X = up<T>(x), EVK = up<T>(evk), P = up<T>(p), PM = up<T>(perm), TW = up<T>(tw), ITW = up<T>(itw),
             MI = up<T>(minv), PERMS = up<T>(perms);
        auto digits = allocate_on_hardware<T>({N, L, D});
        auto dig_ntt = allocate_on_hardware<T>({1, N, D, L});
        auto prod = allocate_on_hardware<T>({D, 2, N, L});
        auto acc = allocate_on_hardware<T>({2, N, L});
        auto coef = allocate_on_hardware<T>({2, N, 1, L});
        auto res = allocate_on_hardware<T>({2, N, L});
        // The API has no transpose: the harness makes the digit-major views by
        // copying host tensors, outside the timed region, exactly as a caller
        // with a moveaxis op would.
        auto t0 = std::chrono::steady_clock::now();
        g_decomposition<T>(X, digits, dnum, base_bits);
        auto t1 = std::chrono::steady_clock::now();
        auto dv = host_to_device<T>(device_to_host<T>(digits).view({1, N, L, D}).permute({0, 1, 3, 2}));
        auto t2 = std::chrono::steady_clock::now();
        ntt<T>(dv, P, PM, TW, nullptr, nullptr, dig_ntt);
        auto t3 = std::chrono::steady_clock::now();
        auto dview = host_to_device<T>(device_to_host<T>(dig_ntt).view({N, D, L}).permute({1, 0, 2}).unsqueeze(1));
        auto t4 = std::chrono::steady_clock::now();
        modmul_ttt<T>(dview, EVK, P, prod);
        axis_modsum<T>(prod, P, acc, 0);
        acc->reshape({2, N, 1, L});
        intt<T>(acc, P, PM, ITW, MI, nullptr, nullptr, coef);
        coef->reshape({2, N, L});
        permute<T>(coef, PERMS, res, 2, 1);
        auto t5 = std::chrono::steady_clock::now();
        down<T>(res, out);
        std::chrono::duration<double> compute = (t1 - t0) + (t3 - t2) + (t5 - t4);
        if (seconds) *seconds = compute.count();
    });
}

}  // extern "C"
