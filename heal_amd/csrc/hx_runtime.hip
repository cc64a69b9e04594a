// hx_runtime.hip — library state, errors, stream-ordered memory, strided copy.
//
// Replaces the storage layer of the reference's DeviceTensor
// (example_impl/device_memory_impl.cpp:14-202), which keeps tensors in host
// malloc memory, with HBM from a stream-ordered caching pool.
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <utility>
#include <vector>

#include "hx_internal.hpp"

namespace hx {

static thread_local std::string t_err;

void raise(int code, const std::string& msg) { throw Error{code, msg}; }

int set_error(int code, const std::string& msg) {
    t_err = msg;
    return code;
}

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory) throw std::bad_alloc();
        raise(HX_ERR_RUNTIME, std::string(what) + ": " + hipGetErrorString(e));
    }
}

size_t dtype_size(int dtype) {
    switch (dtype) {
        case HX_INT32: return 4;
        case HX_INT64: return 8;
        case HX_FLOAT64: return 8;
        default: raise(HX_ERR_INVALID_ARGUMENT, "unknown dtype");
    }
}

void check_view(const hx_view* v, const char* name) {
    if (!v) raise(HX_ERR_INVALID_ARGUMENT, std::string(name) + " pointer must not be null.");
    if (v->ndim < 0 || v->ndim > HX_MAX_DIMS)
        raise(HX_ERR_INVALID_ARGUMENT, std::string(name) + ": ndim must be in [0, 8]");
    dtype_size(v->dtype);
    int64_t n = 1;
    for (int i = 0; i < v->ndim; ++i) {
        if (v->dims[i] < 0) raise(HX_ERR_INVALID_ARGUMENT, std::string(name) + ": negative dimension");
        n *= v->dims[i];
    }
    if (n > 0 && !v->data) raise(HX_ERR_INVALID_ARGUMENT, std::string(name) + ": null data pointer");
}

int64_t view_numel(const hx_view* v) {
    int64_t n = 1;
    for (int i = 0; i < v->ndim; ++i) n *= v->dims[i];
    return n;
}

bool dims_equal(const hx_view* a, const hx_view* b) {
    if (a->ndim != b->ndim) return false;
    for (int i = 0; i < a->ndim; ++i)
        if (a->dims[i] != b->dims[i]) return false;
    return true;
}

void view_byte_range(const hx_view* v, uintptr_t* lo, uintptr_t* hi) {
    if (view_numel(v) == 0) {
        *lo = 1;
        *hi = 0;
        return;
    }
    int64_t mn = 0, mx = 0;
    for (int i = 0; i < v->ndim; ++i) {
        int64_t span = (v->dims[i] - 1) * v->strides[i];
        if (span < 0) mn += span;
        else mx += span;
    }
    size_t es = dtype_size(v->dtype);
    *lo = (uintptr_t)v->data + (intptr_t)(mn * (int64_t)es);
    *hi = (uintptr_t)v->data + (intptr_t)(mx * (int64_t)es) + es - 1;
}

bool views_overlap(const hx_view* a, const hx_view* b) {
    uintptr_t alo, ahi, blo, bhi;
    view_byte_range(a, &alo, &ahi);
    view_byte_range(b, &blo, &bhi);
    if (alo > ahi || blo > bhi) return false;
    return !(ahi < blo || bhi < alo);
}

IterSpace make_iter(int ndim, const int64_t* dims, int nops, const int64_t (*strides)[HX_MAX_DIMS], bool keep_last,
                    int order_op) {
    IterSpace it;
    std::memset(&it, 0, sizeof it);
    int64_t d[HX_MAX_DIMS];
    int64_t s[kMaxOps][HX_MAX_DIMS];
    int n = 0;
    int64_t total = 1;
    for (int i = 0; i < ndim; ++i) total *= dims[i];
    for (int i = 0; i < ndim; ++i) {
        bool is_last = (i == ndim - 1);
        if (dims[i] == 1 && !(keep_last && is_last)) continue;
        d[n] = dims[i];
        for (int o = 0; o < nops; ++o) s[o][n] = strides[o][i];
        ++n;
    }
    // memory order of operand order_op: stable insertion sort of the dims by
    // decreasing |stride| (the last dim stays in place when keep_last)
    {
        const int ns = (keep_last && n > 0) ? n - 1 : n;
        auto key = [&](int i) { return s[order_op][i] < 0 ? -s[order_op][i] : s[order_op][i]; };
        for (int i = 1; i < ns; ++i) {
            for (int j = i; j > 0 && key(j - 1) < key(j); --j) {
                std::swap(d[j - 1], d[j]);
                for (int o = 0; o < nops; ++o) std::swap(s[o][j - 1], s[o][j]);
            }
        }
    }
    // merge dim j into j+1 (j outer) when every operand is contiguous across them
    int m = 0;
    int64_t md[HX_MAX_DIMS];
    int64_t ms[kMaxOps][HX_MAX_DIMS];
    for (int i = 0; i < n; ++i) {
        if (m > 0) {
            bool can = !(keep_last && i == n - 1);
            for (int o = 0; o < nops && can; ++o) can = (ms[o][m - 1] == s[o][i] * d[i]);
            if (can) {
                // outer dim (m-1) absorbs inner dim i: new inner stride is s[o][i]
                md[m - 1] *= d[i];
                for (int o = 0; o < nops; ++o) ms[o][m - 1] = s[o][i];
                continue;
            }
        }
        md[m] = d[i];
        for (int o = 0; o < nops; ++o) ms[o][m] = s[o][i];
        ++m;
    }
    if (m == 0) {
        md[0] = 1;
        for (int o = 0; o < nops; ++o) ms[o][0] = 0;
        m = 1;
    }
    it.ndim = m;
    it.total = total;
    it.use32 = total < (int64_t(1) << 31);
    for (int i = 0; i < m; ++i) {
        it.dims[i] = md[i];
        it.fd[i] = make_fastdiv(it.use32 ? (uint32_t)md[i] : 1u);
        for (int o = 0; o < nops; ++o) it.strides[o][i] = ms[o][i];
    }
    return it;
}

// ---------------------------------------------------------- device status
static std::mutex g_status_mu;
static std::vector<DevStatus*> g_status;

DevStatus* device_status() {
    int dev = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lk(g_status_mu);
    if ((int)g_status.size() <= dev) g_status.resize(dev + 1, nullptr);
    if (!g_status[dev]) {
        DevStatus* p = nullptr;
        check_hip(hipMalloc((void**)&p, sizeof(DevStatus)), "hipMalloc(status)");
        check_hip(hipMemset(p, 0, sizeof(DevStatus)), "hipMemset(status)");
        check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
        g_status[dev] = p;
        // keep freed pool memory cached in the process instead of returning it
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
            uint64_t thr = ~0ull;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
        }
    }
    return g_status[dev];
}

// ------------------------------------------------------------ fill / copy
__global__ void k_fill_zero(uint64_t* p8, size_t n8, unsigned char* tail, size_t ntail) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t j = i; j < n8; j += stride) p8[j] = 0;
    if (i < ntail) tail[i] = 0;
}

template <typename E>
__global__ void k_strided_copy(IterSpace it, const E* __restrict__ src, E* __restrict__ dst) {
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < it.total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        int64_t off[2], last;
        iter_offsets<2>(it, idx, off, &last);
        dst[off[1]] = src[off[0]];
    }
}

void strided_copy(const hx_view* src, const hx_view* dst, hipStream_t st) {
    int64_t strides[kMaxOps][HX_MAX_DIMS] = {};
    for (int i = 0; i < src->ndim; ++i) {
        strides[0][i] = src->strides[i];
        strides[1][i] = dst->strides[i];
    }
    IterSpace it = make_iter(src->ndim, src->dims, 2, strides, false, 1);  // destination order
    if (it.total == 0) return;
    unsigned g = grid_for(it.total, 256, 8192);
    if (dtype_size(src->dtype) == 4)
        hipLaunchKernelGGL(k_strided_copy<uint32_t>, dim3(g), dim3(256), 0, st, it, (const uint32_t*)src->data,
                           (uint32_t*)dst->data);
    else
        hipLaunchKernelGGL(k_strided_copy<uint64_t>, dim3(g), dim3(256), 0, st, it, (const uint64_t*)src->data,
                           (uint64_t*)dst->data);
    check_hip(hipGetLastError(), "strided_copy launch");
}

void fill_zero(void* ptr, size_t bytes, hipStream_t st) {
    if (!bytes) return;
    uintptr_t p = (uintptr_t)ptr;
    if (p % 8) raise(HX_ERR_INVALID_ARGUMENT, "hx_memset_zero: pointer must be 8-byte aligned");
    size_t n8 = bytes / 8, ntail = bytes % 8;
    unsigned g = grid_for((int64_t)(n8 ? n8 : 1), 256, 8192);
    hipLaunchKernelGGL(k_fill_zero, dim3(g), dim3(256), 0, st, (uint64_t*)ptr, n8, (unsigned char*)ptr + n8 * 8,
                       ntail);
    check_hip(hipGetLastError(), "fill launch");
}

}  // namespace hx

using namespace hx;

extern "C" {

int hx_abi_version(void) { return HX_ABI_VERSION; }

const char* hx_last_error(void) { return t_err.c_str(); }

int hx_status_poll(hx_stream stream, hx_device_status* out) {
    HX_API_BEGIN
    DevStatus* st = device_status();
    check_hip(hipStreamSynchronize(as_stream(stream)), "hipStreamSynchronize");
    DevStatus h;
    check_hip(hipMemcpy(&h, st, sizeof h, hipMemcpyDeviceToHost), "hipMemcpy(status)");
    check_hip(hipMemset(st, 0, sizeof h), "hipMemset(status)");
    check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
    if (out) {
        out->gdecomp_overflow = h.gdecomp_overflow;
        out->index_out_of_range = h.index_out_of_range;
    }
    HX_API_END
}

int hx_malloc(void** out, size_t bytes, hx_stream stream) {
    HX_API_BEGIN
    if (!out) raise(HX_ERR_INVALID_ARGUMENT, "out pointer must not be null.");
    *out = nullptr;
    device_status();  // initialises the pool policy for this device
    if (bytes == 0) return HX_OK;
    check_hip(hipMallocAsync(out, bytes, as_stream(stream)), "hipMallocAsync");
    HX_API_END
}

int hx_free(void* ptr, hx_stream stream) {
    HX_API_BEGIN
    if (ptr) check_hip(hipFreeAsync(ptr, as_stream(stream)), "hipFreeAsync");
    HX_API_END
}

int hx_memset_zero(void* ptr, size_t bytes, hx_stream stream) {
    HX_API_BEGIN
    fill_zero(ptr, bytes, as_stream(stream));
    HX_API_END
}

int hx_memcpy_h2d(void* dst, const void* src, size_t bytes, hx_stream stream) {
    HX_API_BEGIN
    if (!bytes) return HX_OK;
    check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)), "hipMemcpyAsync(H2D)");
    check_hip(hipStreamSynchronize(as_stream(stream)), "hipStreamSynchronize");
    HX_API_END
}

int hx_memcpy_d2h(void* dst, const void* src, size_t bytes, hx_stream stream) {
    HX_API_BEGIN
    if (!bytes) return HX_OK;
    check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)), "hipMemcpyAsync(D2H)");
    check_hip(hipStreamSynchronize(as_stream(stream)), "hipStreamSynchronize");
    HX_API_END
}

int hx_memcpy_d2d(void* dst, const void* src, size_t bytes, hx_stream stream) {
    HX_API_BEGIN
    if (!bytes) return HX_OK;
    check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)), "hipMemcpyAsync(D2D)");
    HX_API_END
}

int hx_stream_sync(hx_stream stream) {
    HX_API_BEGIN
    check_hip(hipStreamSynchronize(as_stream(stream)), "hipStreamSynchronize");
    HX_API_END
}

int hx_strided_copy(const hx_view* src, const hx_view* dst, hx_stream stream) {
    HX_API_BEGIN
    check_view(src, "src");
    check_view(dst, "dst");
    if (!dims_equal(src, dst)) raise(HX_ERR_INVALID_ARGUMENT, "src and dst must have the same shape");
    if (dtype_size(src->dtype) != dtype_size(dst->dtype)) raise(HX_ERR_INVALID_ARGUMENT, "src and dst dtypes differ");
    strided_copy(src, dst, as_stream(stream));
    HX_API_END
}

}  // extern "C"
