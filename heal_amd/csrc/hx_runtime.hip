// hx_runtime.hip — library state, errors, stream-ordered memory, strided copy.
//
// Replaces the storage layer of the reference's DeviceTensor
// (example_impl/device_memory_impl.cpp:14-202), which keeps tensors in host
// malloc memory, with HBM from a stream-ordered caching pool.
#include <algorithm>
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
        // once per device: the block is zeroed before any kernel can see it
        DevStatus* p = nullptr;
        check_hip(hipMalloc((void**)&p, sizeof(DevStatus)), "hipMalloc(status)");
        check_hip(hipMemcpy(p, std::vector<char>(sizeof(DevStatus), 0).data(), sizeof(DevStatus),
                            hipMemcpyHostToDevice),
                  "hipMemcpy(status)");
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

__global__ void k_status_take(unsigned long long* st, unsigned long long* out, int n) {
    const int i = threadIdx.x;
    if (i < n) out[i] = atomicExch(&st[i], 0ull);
}

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

void fill_u32(uint32_t* ptr, int64_t n, uint32_t value, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, ptr, n, value);
    check_hip(hipGetLastError(), "fill launch");
}

// --------------------------------------------------- side and reclaim streams
// Two library streams per device (default priority, non-blocking): the side
// stream runs host-synchronising work while a caller's stream is being
// captured, the reclaim stream orders frees after the work that used the
// memory.  They live until the process ends.
struct DevStreams {
    hipStream_t side = nullptr;
    hipStream_t reclaim = nullptr;
};
static std::mutex g_streams_mu;
static std::vector<DevStreams> g_streams;

static DevStreams& dev_streams(int dev) {
    std::lock_guard<std::mutex> lk(g_streams_mu);
    if ((int)g_streams.size() <= dev) g_streams.resize(dev + 1);
    DevStreams& s = g_streams[dev];
    if (!s.side) {
        int cur = 0;
        check_hip(hipGetDevice(&cur), "hipGetDevice");
        if (cur != dev) check_hip(hipSetDevice(dev), "hipSetDevice");
        hipError_t e1 = hipStreamCreateWithFlags(&s.side, hipStreamNonBlocking);
        hipError_t e2 = hipStreamCreateWithFlags(&s.reclaim, hipStreamNonBlocking);
        if (cur != dev) (void)hipSetDevice(cur);
        check_hip(e1, "hipStreamCreateWithFlags");
        check_hip(e2, "hipStreamCreateWithFlags");
    }
    return s;
}

bool stream_capturing(hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    check_hip(hipStreamIsCapturing(st, &cs), "hipStreamIsCapturing");
    return cs != hipStreamCaptureStatusNone;
}

SyncScope::SyncScope(hipStream_t st) : st_(st), escaped_(false), mode_(0) {
    if (!stream_capturing(st)) return;
    int dev = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    hipStreamCaptureMode m = hipStreamCaptureModeRelaxed;
    check_hip(hipThreadExchangeStreamCaptureMode(&m), "hipThreadExchangeStreamCaptureMode");
    mode_ = (int)m;
    escaped_ = true;
    st_ = dev_streams(dev).side;
}

SyncScope::~SyncScope() {
    if (!escaped_) return;
    hipStreamCaptureMode m = (hipStreamCaptureMode)mode_;
    (void)hipThreadExchangeStreamCaptureMode(&m);
}

void SyncScope::sync() { check_hip(hipStreamSynchronize(st_), "hipStreamSynchronize"); }

void release_after(int device, std::vector<hipEvent_t>& events, const std::vector<void*>& ptrs) {
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) cur = device;
    if (cur != device) (void)hipSetDevice(device);
    hipStream_t rs = nullptr;
    try {
        rs = dev_streams(device).reclaim;
    } catch (...) {
        rs = nullptr;
    }
    for (hipEvent_t ev : events) {
        if (rs) (void)hipStreamWaitEvent(rs, ev, 0);
        (void)hipEventDestroy(ev);
    }
    events.clear();
    for (void* p : ptrs)
        if (p) {
            if (rs) (void)hipFreeAsync(p, rs);
            else (void)hipFree(p);
        }
    if (cur != device) (void)hipSetDevice(cur);
}

// ----------------------------------------------------------------- workspaces
// A workspace allocated for a capture is made outside the graph: on the side
// stream with the capture mode relaxed, flags zeroed there, then waited for.
static Workspaces::Ws alloc_ws(hipStream_t st, bool capturing, size_t flag_bytes, size_t data_bytes) {
    Workspaces::Ws w;
    SyncScope sc(st);
    hipStream_t s = sc.stream();
    const size_t fb = (flag_bytes + 255) / 256 * 256;
    if (fb + data_bytes == 0) return w;
    void* p = nullptr;
    check_hip(hipMallocAsync(&p, fb + data_bytes, s), "hipMallocAsync(workspace)");
    if (fb) fill_zero(p, fb, s);
    if (capturing) sc.sync();
    w.flags = fb ? p : nullptr;
    w.flag_bytes = fb;
    w.data = data_bytes ? (char*)p + fb : nullptr;
    w.data_bytes = data_bytes;
    return w;
}

static void* ws_base(const Workspaces::Ws& w) { return w.flags ? w.flags : w.data; }

Workspaces::Ws Workspaces::acquire(hipStream_t st, size_t flag_bytes, size_t data_bytes) {
    if (stream_capturing(st)) {
        // the graph keeps this workspace; it is released with the plan
        Ws w = alloc_ws(st, true, flag_bytes, data_bytes);
        captured.push_back(w);
        return w;
    }
    for (Eager& e : eager) {
        if (e.s != st) continue;
        if (e.ws.flag_bytes >= flag_bytes && e.ws.data_bytes >= data_bytes) return e.ws;
        // grown in this stream's order: earlier work on st still reads the old one
        const size_t fb = std::max(flag_bytes, e.ws.flag_bytes), db = std::max(data_bytes, e.ws.data_bytes);
        if (void* b = ws_base(e.ws)) check_hip(hipFreeAsync(b, st), "hipFreeAsync(workspace)");
        e.ws = Ws();
        e.ws = alloc_ws(st, false, fb, db);
        return e.ws;
    }
    Eager e;
    e.s = st;
    e.ev = nullptr;
    check_hip(hipEventCreateWithFlags(&e.ev, hipEventDisableTiming), "hipEventCreate");
    try {
        e.ws = alloc_ws(st, false, flag_bytes, data_bytes);
    } catch (...) {
        (void)hipEventDestroy(e.ev);
        throw;
    }
    eager.push_back(e);
    return e.ws;
}

void Workspaces::note(hipStream_t st) {
    if (stream_capturing(st)) return;
    for (Eager& e : eager)
        if (e.s == st) {
            check_hip(hipEventRecord(e.ev, st), "hipEventRecord");
            return;
        }
    // a stream that needed no workspace still uses the plan's tables
    Eager e;
    e.s = st;
    check_hip(hipEventCreateWithFlags(&e.ev, hipEventDisableTiming), "hipEventCreate");
    eager.push_back(e);
    check_hip(hipEventRecord(e.ev, st), "hipEventRecord");
}

void Workspaces::release(std::vector<hipEvent_t>& events, std::vector<void*>& ptrs) {
    for (Eager& e : eager) {
        events.push_back(e.ev);
        ptrs.push_back(ws_base(e.ws));
    }
    for (Ws& w : captured) ptrs.push_back(ws_base(w));
    eager.clear();
    captured.clear();
}

}  // namespace hx

using namespace hx;

extern "C" {

int hx_abi_version(void) { return HX_ABI_VERSION; }

const char* hx_last_error(void) { return t_err.c_str(); }

int hx_status_poll_log(hx_stream stream, hx_device_status* out, hx_gdecomp_log* log) {
    HX_API_BEGIN
    // in the stream's order: counters and log are taken (read and cleared
    // atomically, word by word) after this stream's earlier work; counts of
    // work still running on other streams stay for the next poll
    static_assert(kGdLog == HX_GDECOMP_LOG_MAX, "log size");
    DevStatus* st = device_status();
    hipStream_t s = as_stream(stream);
    const int n = (int)(sizeof(DevStatus) / sizeof(unsigned long long));
    static_assert(sizeof(DevStatus) / sizeof(unsigned long long) <= 256, "one workgroup takes the status");
    void* tmp = nullptr;
    check_hip(hipMallocAsync(&tmp, sizeof(DevStatus), s), "hipMallocAsync(status)");
    DevStatus h;
    hipLaunchKernelGGL(k_status_take, dim3(1), dim3(256), 0, s, (unsigned long long*)st, (unsigned long long*)tmp, n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&h, tmp, sizeof h, hipMemcpyDeviceToHost, s);
    (void)hipFreeAsync(tmp, s);
    check_hip(e, "status take");
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (out) {
        out->gdecomp_overflow = h.gdecomp_overflow;
        out->index_out_of_range = h.index_out_of_range;
    }
    if (log) {
        const uint64_t k = std::min<uint64_t>(h.gd_log_count, (uint64_t)kGdLog);
        log->count = k;
        for (uint64_t i = 0; i < k; ++i) {
            log->tag[i] = h.gd_log[2 * i];
            log->index[i] = h.gd_log[2 * i + 1];
        }
    }
    HX_API_END
}

int hx_status_poll(hx_stream stream, hx_device_status* out) { return hx_status_poll_log(stream, out, nullptr); }

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

int hx_stream_capturing(hx_stream stream, int* capturing) {
    HX_API_BEGIN
    if (!capturing) raise(HX_ERR_INVALID_ARGUMENT, "capturing pointer must not be null.");
    *capturing = stream_capturing(as_stream(stream)) ? 1 : 0;
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
