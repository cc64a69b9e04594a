// lattica_hw_api.cpp — the HEAL hardware API templates on the MI355X C ABI.
//
// Memory and views follow example_impl/device_memory_impl.cpp,
// memory_virtual_ops_impl.cpp and contiguous_impl.cpp of the reference
// (semantics restated in each function's comment); the compute ops validate
// nothing beyond argument presence here and pass views to include/heal_amd.h,
// which applies the reference's checks.  NTT tables and Galois tables are
// prepared once per (buffer id, write version, view) and reused.
#include "include/lattica_hw_api.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <functional>
#include <iostream>
#include <list>
#include <map>
#include <mutex>
#include <new>
#include <numeric>
#include <type_traits>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <typeinfo>
#include <vector>

#include "../../include/heal_amd.h"

namespace lattica_hw_api {
namespace detail {

static thread_local void* t_stream = nullptr;
static std::atomic<uint64_t> g_next_id{1};

[[noreturn]] void throw_status(int code) {
    std::string msg = hx_last_error();
    switch (code) {
        case HX_ERR_INVALID_ARGUMENT: throw std::invalid_argument(msg);
        case HX_ERR_OUT_OF_RANGE: throw std::out_of_range(msg);
        case HX_ERR_BAD_ALLOC: throw std::bad_alloc();
        default: throw std::runtime_error(msg);
    }
}

inline void check(int code) {
    if (code != HX_OK) throw_status(code);
}

Buffer::~Buffer() {
    if (ptr) hx_free(ptr, t_stream);
}

std::shared_ptr<Buffer> new_buffer(size_t bytes, bool zero) {
    auto b = std::make_shared<Buffer>();
    b->bytes = bytes;
    b->id = g_next_id++;
    if (bytes) {
        check(hx_malloc(&b->ptr, bytes, t_stream));
        if (zero) check(hx_memset_zero(b->ptr, bytes, t_stream));
    }
    return b;
}

template <typename T>
constexpr int dtype_code() {
    return std::is_same<T, int32_t>::value ? HX_INT32 : (std::is_same<T, int64_t>::value ? HX_INT64 : HX_FLOAT64);
}

template <typename T>
hx_view view_of(const DeviceTensor<T>& t) {
    if (t.dims.size() > HX_MAX_DIMS) throw std::invalid_argument("tensors of more than 8 dims are not supported");
    hx_view v;
    std::memset(&v, 0, sizeof v);
    v.data = t.data_ptr();
    v.dtype = dtype_code<T>();
    v.ndim = (int32_t)t.dims.size();
    for (size_t i = 0; i < t.dims.size(); ++i) {
        v.dims[i] = t.dims[i];
        v.strides[i] = t.strides[i];
    }
    return v;
}

template <typename T>
void bump(const std::shared_ptr<DeviceTensor<T>>& t) {
    if (t && t->buffer) ++t->buffer->version;
}

int64_t span_elems(const std::vector<int64_t>& dims, const std::vector<int64_t>& strides) {
    int64_t s = 1;
    for (size_t i = 0; i < dims.size(); ++i) {
        if (dims[i] == 0) return 0;
        s += (dims[i] - 1) * strides[i];
    }
    return s;
}

std::vector<int64_t> contiguous_strides(const std::vector<int64_t>& dims) {
    std::vector<int64_t> s(dims.size());
    int64_t st = 1;
    for (int64_t i = (int64_t)dims.size() - 1; i >= 0; --i) {
        s[i] = st;
        st *= dims[i];
    }
    return s;
}

// Deferred reporting of data-dependent conditions: the elements of a
// g_decomposition whose values were not exhausted are reported with the
// reference's stderr warning (example_impl/g_decomposition_impl.cpp:58-65)
// when the stream is next synchronised.  Each call carries a tag into the
// device's overflow log; the tag names the call's shape, base_bits and power.
struct GdCall {
    std::vector<int64_t> dims;
    size_t power = 0, base_bits = 0;
};
struct PendingWarn {
    std::mutex mu;
    uint64_t next_tag = 0;
    std::map<uint64_t, GdCall> calls;  // recent calls by tag (the oldest are dropped beyond kKeep)
    static constexpr size_t kKeep = 4096;
};
static PendingWarn g_warn;

static void print_gdecomp_warning(const GdCall& c, uint64_t index) {
    std::vector<int64_t> coord(c.dims.size());
    for (size_t i = c.dims.size(); i-- > 0;) {
        const int64_t d = c.dims[i] > 0 ? c.dims[i] : 1;
        coord[i] = (int64_t)(index % (uint64_t)d);
        index /= (uint64_t)d;
    }
    std::cerr << "Warning: value at ";
    for (auto x : coord) std::cerr << x << " ";
    std::cerr << "exceeds capacity with base_bits=" << c.base_bits << " and power=" << c.power << "\n";
}

void sync_and_report() {
    hx_device_status st;
    hx_gdecomp_log log;
    check(hx_status_poll_log(t_stream, &st, &log));
    if (!st.gdecomp_overflow) return;
    std::lock_guard<std::mutex> lk(g_warn.mu);
    std::vector<std::pair<uint64_t, uint64_t>> entries;
    for (uint64_t i = 0; i < log.count; ++i) entries.emplace_back(log.tag[i], log.index[i]);
    std::sort(entries.begin(), entries.end());
    uint64_t printed = 0;
    for (const auto& e : entries) {
        auto it = g_warn.calls.find(e.first);
        if (it == g_warn.calls.end()) continue;
        print_gdecomp_warning(it->second, e.second);
        ++printed;
    }
    if (st.gdecomp_overflow > printed)
        std::cerr << "Warning: " << (st.gdecomp_overflow - printed)
                  << (printed ? " more" : "") << " value(s) exceed capacity in g_decomposition\n";
}

// ---- caches of prepared device tables, keyed by buffer identity and version
struct TensorKey {
    uint64_t id = 0, version = 0;
    int64_t offset = 0;
    std::vector<int64_t> dims, strides;
    bool operator<(const TensorKey& o) const {
        return std::tie(id, version, offset, dims, strides) < std::tie(o.id, o.version, o.offset, o.dims, o.strides);
    }
    bool operator==(const TensorKey& o) const {
        return std::tie(id, version, offset, dims, strides) == std::tie(o.id, o.version, o.offset, o.dims, o.strides);
    }
};

template <typename T>
TensorKey key_of(const std::shared_ptr<DeviceTensor<T>>& t) {
    TensorKey k;
    if (!t) return k;
    k.id = t->buffer ? t->buffer->id : 0;
    k.version = t->buffer ? t->buffer->version : 0;
    k.offset = t->offset;
    k.dims = t->dims;
    k.strides = t->strides;
    return k;
}

struct PlanKey {
    int inverse, dtype;
    TensorKey p, perm, tw, minv;
    bool operator<(const PlanKey& o) const {
        return std::tie(inverse, dtype, p, perm, tw, minv) < std::tie(o.inverse, o.dtype, o.p, o.perm, o.tw, o.minv);
    }
    bool operator==(const PlanKey& o) const {
        return std::tie(inverse, dtype, p, perm, tw, minv) == std::tie(o.inverse, o.dtype, o.p, o.perm, o.tw, o.minv);
    }
};

struct PlanHandle {
    hx_ntt_plan* plan = nullptr;
    ~PlanHandle() {
        if (plan) hx_ntt_plan_destroy(plan);
    }
};

constexpr size_t kMaxPlans = 64;
static std::mutex g_cache_mu;
static std::map<PlanKey, std::shared_ptr<PlanHandle>> g_plans;
static std::list<PlanKey> g_plan_lru;
// validated perms tables -> every row equals row 0 (the gather then reads row 0 only)
static std::map<TensorKey, bool> g_valid_perms;

template <typename T>
std::shared_ptr<PlanHandle> get_plan(int inverse, const std::shared_ptr<DeviceTensor<T>>& p,
                                     const std::shared_ptr<DeviceTensor<T>>& perm,
                                     const std::shared_ptr<DeviceTensor<T>>& tw,
                                     const std::shared_ptr<DeviceTensor<T>>& minv) {
    PlanKey k{inverse, dtype_code<T>(), key_of(p), key_of(perm), key_of(tw), key_of(minv)};
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto it = g_plans.find(k);
        if (it != g_plans.end()) {
            g_plan_lru.remove(k);
            g_plan_lru.push_front(k);
            return it->second;
        }
    }
    hx_view pv = view_of(*p), permv = view_of(*perm), twv = view_of(*tw);
    hx_view mv;
    if (minv) mv = view_of(*minv);
    auto h = std::make_shared<PlanHandle>();
    check(hx_ntt_plan_create(&h->plan, inverse, &pv, &permv, &twv, minv ? &mv : nullptr, t_stream));
    std::lock_guard<std::mutex> lk(g_cache_mu);
    g_plans[k] = h;
    g_plan_lru.push_front(k);
    while (g_plan_lru.size() > kMaxPlans) {
        g_plans.erase(g_plan_lru.back());
        g_plan_lru.pop_back();
    }
    return h;
}

template <typename T>
void require(const std::shared_ptr<DeviceTensor<T>>& t, const char* name) {
    if (!t) throw std::invalid_argument(std::string(name) + " pointer must not be null.");
}

// ---- plans held by captured graphs
static thread_local std::shared_ptr<CapturePins> t_pins;
static std::mutex g_pin_mu;
static std::vector<std::shared_ptr<PlanHandle>> g_forever_pins;  // captures outside a record
static std::atomic<int64_t> g_pinned{0};

bool stream_capturing() {
    int c = 0;
    check(hx_stream_capturing(t_stream, &c));
    return c != 0;
}

}  // namespace detail

struct CapturePins {
    std::vector<std::shared_ptr<detail::PlanHandle>> plans;
    ~CapturePins() { detail::g_pinned -= (int64_t)plans.size(); }
};

namespace detail {

// A plan used by an op while the stream is being captured stays alive with the graph.
void pin_if_capturing(const std::shared_ptr<PlanHandle>& plan) {
    if (!stream_capturing()) return;
    ++g_pinned;
    if (t_pins) {
        t_pins->plans.push_back(plan);
    } else {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        g_forever_pins.push_back(plan);
    }
}

}  // namespace detail

using namespace detail;

void set_stream(void* stream) { t_stream = stream; }
void* get_stream() { return t_stream; }
void synchronize() { sync_and_report(); }
size_t plan_cache_size() {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    return g_plans.size() + g_valid_perms.size();
}
void capture_begin() { t_pins = std::make_shared<CapturePins>(); }
std::shared_ptr<CapturePins> capture_end() {
    auto p = t_pins ? t_pins : std::make_shared<CapturePins>();
    t_pins.reset();
    return p;
}
size_t captured_plan_count() { return (size_t)g_pinned.load(); }

// ------------------------------------------------------------------ memory

// allocate_on_hardware (example_impl/device_memory_impl.cpp:168-179): zeroed, C-contiguous
template <typename T>
std::shared_ptr<DeviceTensor<T>> allocate_on_hardware(const std::vector<int64_t>& dims) {
    for (int64_t d : dims)
        if (d < 0) throw std::invalid_argument("dimensions must be non-negative");
    auto t = std::make_shared<DeviceTensor<T>>();
    t->dims = dims;
    t->strides = contiguous_strides(dims);
    int64_t n = std::accumulate(dims.begin(), dims.end(), int64_t(1), std::multiplies<int64_t>());
    t->buffer = new_buffer((size_t)n * sizeof(T), true);
    return t;
}

// host_to_device (example_impl/device_memory_impl.cpp:181-190): dtype must be T;
// the tensor's strided span is copied and its strides are kept
template <typename T>
std::shared_ptr<DeviceTensor<T>> host_to_device(const torch::Tensor& tensor) {
    if (tensor.scalar_type() != torch::CppTypeToScalarType<T>())
        throw std::runtime_error("Tensor dtype does not match template parameter T.");
    auto t = std::make_shared<DeviceTensor<T>>();
    t->dims.assign(tensor.sizes().begin(), tensor.sizes().end());
    t->strides.assign(tensor.strides().begin(), tensor.strides().end());
    int64_t span = span_elems(t->dims, t->strides);
    t->buffer = new_buffer((size_t)span * sizeof(T), false);
    if (span) {
        const void* src = tensor.data_ptr();
        if (tensor.is_cuda())
            check(hx_memcpy_d2d(t->buffer->ptr, src, (size_t)span * sizeof(T), t_stream));
        else
            check(hx_memcpy_h2d(t->buffer->ptr, src, (size_t)span * sizeof(T), t_stream));
    }
    return t;
}

// device_to_host (example_impl/device_memory_impl.cpp:192-202): a host tensor
// with the device tensor's dims and strides, cloned
template <typename T>
torch::Tensor device_to_host(const std::shared_ptr<DeviceTensor<T>>& memory) {
    require(memory, "memory");
    sync_and_report();
    int64_t span = span_elems(memory->dims, memory->strides);
    std::vector<T> host((size_t)std::max<int64_t>(span, 1));
    if (span) check(hx_memcpy_d2h(host.data(), memory->data_ptr(), (size_t)span * sizeof(T), t_stream));
    auto opts = torch::TensorOptions().dtype(torch::CppTypeToScalarType<T>());
    return torch::from_blob(host.data(), memory->dims, memory->strides, opts).clone();
}

}  // namespace lattica_hw_api

using namespace lattica_hw_api::detail;

template <typename T>
bool DeviceTensor<T>::is_contiguous() const {
    int64_t expected = 1;
    for (int64_t i = (int64_t)dims.size() - 1; i >= 0; --i) {
        if (dims[i] == 1) continue;
        if (strides[i] != expected) return false;
        expected *= dims[i];
    }
    return true;
}

template <typename T>
void DeviceTensor<T>::reshape(const std::vector<int64_t>& new_dims) {
    int64_t new_total = std::accumulate(new_dims.begin(), new_dims.end(), int64_t(1), std::multiplies<int64_t>());
    int64_t current = 1;
    for (size_t i = 0; i < dims.size(); ++i)
        if (strides[i] != 0) current *= dims[i];
    if (new_total != current)
        throw std::invalid_argument("Total size of new shape must match number of elements (excluding broadcasted dims).");
    dims = new_dims;
    strides = contiguous_strides(new_dims);
}

template <typename T>
void DeviceTensor<T>::print_metadata() const {
    std::cout << "DeviceTensor<" << typeid(T).name() << "> Shape: [";
    for (auto d : dims) std::cout << d << " ";
    std::cout << "]  Strides: [";
    for (auto s : strides) std::cout << s << " ";
    std::cout << "]\n\n";
}

template <typename T>
void DeviceTensor<T>::print() const {
    sync_and_report();
    int64_t span = span_elems(dims, strides);
    std::vector<T> host((size_t)std::max<int64_t>(span, 1));
    if (span) check(hx_memcpy_d2h(host.data(), data_ptr(), (size_t)span * sizeof(T), t_stream));
    std::cout << "DeviceTensor<" << typeid(T).name() << "> Shape: [";
    for (auto d : dims) std::cout << d << " ";
    std::cout << "]  Strides: [";
    for (auto s : strides) std::cout << s << " ";
    std::cout << "]\nData: ";
    std::vector<int64_t> idx(dims.size(), 0);
    std::function<void(size_t)> rec = [&](size_t d) {
        if (d == dims.size()) {
            int64_t off = 0;
            for (size_t i = 0; i < dims.size(); ++i) off += idx[i] * strides[i];
            std::cout << host[(size_t)off];
            return;
        }
        std::cout << "{";
        for (int64_t i = 0; i < dims[d]; ++i) {
            idx[d] = i;
            if (i) std::cout << ", ";
            rec(d + 1);
        }
        std::cout << "}";
    };
    rec(0);
    std::cout << "\n\n";
}

template struct DeviceTensor<int32_t>;
template struct DeviceTensor<int64_t>;
template struct DeviceTensor<double>;

namespace lattica_hw_api {

// ------------------------------------------------------------------ views

// expand (example_impl/memory_virtual_ops_impl.cpp:7-40): in place, size-1 axis -> repeats, stride 0
template <typename T>
std::shared_ptr<DeviceTensor<T>> expand(const std::shared_ptr<DeviceTensor<T>>& a, int64_t axis, int64_t repeats) {
    require(a, "tensor");
    if (repeats <= 0) throw std::invalid_argument("Repeat count must be positive.");
    const int64_t nd = (int64_t)a->dims.size();
    if (axis < 0) axis += nd;
    if (axis < 0 || axis >= nd) throw std::invalid_argument("Invalid repeat dimension.");
    if (a->dims[axis] != 1) throw std::invalid_argument("Can only expand a dimension of size 1.");
    a->dims[axis] *= repeats;
    a->strides[axis] = 0;
    return a;
}

// squeeze (memory_virtual_ops_impl.cpp:42-64): remove a size-1 axis in place
template <typename T>
std::shared_ptr<DeviceTensor<T>> squeeze(const std::shared_ptr<DeviceTensor<T>>& a, int64_t axis) {
    require(a, "tensor");
    const int64_t nd = (int64_t)a->dims.size();
    if (axis < 0) axis += nd;
    if (axis < 0 || axis >= nd) throw std::invalid_argument("Invalid squeeze dimension.");
    if (a->dims[axis] != 1) throw std::invalid_argument("Can only squeeze dimensions of size 1.");
    a->dims.erase(a->dims.begin() + axis);
    a->strides.erase(a->strides.begin() + axis);
    return a;
}

// unsqueeze (memory_virtual_ops_impl.cpp:66-86): insert a size-1 axis with stride 0, in place
template <typename T>
std::shared_ptr<DeviceTensor<T>> unsqueeze(const std::shared_ptr<DeviceTensor<T>>& a, int64_t axis) {
    require(a, "tensor");
    const int64_t nd = (int64_t)a->dims.size();
    if (axis < 0) axis += nd + 1;
    if (axis < 0 || axis > nd) throw std::invalid_argument("Invalid unsqueeze dimension.");
    a->dims.insert(a->dims.begin() + axis, 1);
    a->strides.insert(a->strides.begin() + axis, 0);
    return a;
}

// make_contiguous (example_impl/contiguous_impl.cpp:9-52): the same tensor if
// contiguous; otherwise gather into a new buffer and rebind in place
template <typename T>
std::shared_ptr<DeviceTensor<T>> make_contiguous(const std::shared_ptr<DeviceTensor<T>>& t) {
    require(t, "tensor");
    if (t->is_contiguous()) return t;
    int64_t n = std::accumulate(t->dims.begin(), t->dims.end(), int64_t(1), std::multiplies<int64_t>());
    auto buf = new_buffer((size_t)n * sizeof(T), false);
    hx_view src = view_of(*t);
    DeviceTensor<T> dst_t;
    dst_t.dims = t->dims;
    dst_t.strides = contiguous_strides(t->dims);
    dst_t.buffer = buf;
    hx_view dst = view_of(dst_t);
    check(hx_strided_copy(&src, &dst, t_stream));
    t->buffer = buf;
    t->offset = 0;
    t->strides = dst_t.strides;
    return t;
}

// ------------------------------------------------------------------ compute

#define HX_V(t) (view_of(*(t)))

template <typename T>
void modmul_ttt(const TensorPtr<T>& a, const TensorPtr<T>& b, const TensorPtr<T>& p, TensorPtr<T>& result) {
    require(a, "a"), require(b, "b"), require(p, "p"), require(result, "result");
    hx_view av = HX_V(a), bv = HX_V(b), pv = HX_V(p), rv = HX_V(result);
    check(hx_modmul_ttt(&av, &bv, &pv, &rv, t_stream));
    bump(result);
}
template <typename T>
void modmul_ttc(const TensorPtr<T>& a, const TensorPtr<T>& b, T p, TensorPtr<T>& result) {
    require(a, "a"), require(b, "b"), require(result, "result");
    hx_view av = HX_V(a), bv = HX_V(b), rv = HX_V(result);
    check(hx_modmul_ttc(&av, &bv, (int64_t)p, &rv, t_stream));
    bump(result);
}
template <typename T>
void modmul_tct(const TensorPtr<T>& a, T b, const TensorPtr<T>& p, TensorPtr<T>& result) {
    require(a, "a"), require(p, "p"), require(result, "result");
    hx_view av = HX_V(a), pv = HX_V(p), rv = HX_V(result);
    check(hx_modmul_tct(&av, (int64_t)b, &pv, &rv, t_stream));
    bump(result);
}
template <typename T>
void modmul_tcc(const TensorPtr<T>& a, T b, T p, TensorPtr<T>& result) {
    require(a, "a"), require(result, "result");
    hx_view av = HX_V(a), rv = HX_V(result);
    check(hx_modmul_tcc(&av, (int64_t)b, (int64_t)p, &rv, t_stream));
    bump(result);
}
template <typename T>
void modsum_ttt(const TensorPtr<T>& a, const TensorPtr<T>& b, const TensorPtr<T>& p, TensorPtr<T>& result) {
    require(a, "a"), require(b, "b"), require(p, "p"), require(result, "result");
    hx_view av = HX_V(a), bv = HX_V(b), pv = HX_V(p), rv = HX_V(result);
    check(hx_modsum_ttt(&av, &bv, &pv, &rv, t_stream));
    bump(result);
}
template <typename T>
void modsum_ttc(const TensorPtr<T>& a, const TensorPtr<T>& b, T p, TensorPtr<T>& result) {
    require(a, "a"), require(b, "b"), require(result, "result");
    hx_view av = HX_V(a), bv = HX_V(b), rv = HX_V(result);
    check(hx_modsum_ttc(&av, &bv, (int64_t)p, &rv, t_stream));
    bump(result);
}
template <typename T>
void modsum_tct(const TensorPtr<T>& a, T b, const TensorPtr<T>& p, TensorPtr<T>& result) {
    require(a, "a"), require(p, "p"), require(result, "result");
    hx_view av = HX_V(a), pv = HX_V(p), rv = HX_V(result);
    check(hx_modsum_tct(&av, (int64_t)b, &pv, &rv, t_stream));
    bump(result);
}
template <typename T>
void modsum_tcc(const TensorPtr<T>& a, T b, T p, TensorPtr<T>& result) {
    require(a, "a"), require(result, "result");
    hx_view av = HX_V(a), rv = HX_V(result);
    check(hx_modsum_tcc(&av, (int64_t)b, (int64_t)p, &rv, t_stream));
    bump(result);
}

// mod_* (example_impl/modop_impl.cpp:183-235): null tensors are invalid_argument
template <typename T>
void mod_tt(const TensorPtr<T>& a, const TensorPtr<T>& b, TensorPtr<T>& result) {
    require(a, "a"), require(b, "b"), require(result, "result");
    hx_view av = HX_V(a), bv = HX_V(b), rv = HX_V(result);
    check(hx_mod_tt(&av, &bv, &rv, t_stream));
    bump(result);
}
template <typename T>
void mod_tc(const TensorPtr<T>& a, int64_t b, TensorPtr<T>& result) {
    require(a, "a"), require(result, "result");
    hx_view av = HX_V(a), rv = HX_V(result);
    check(hx_mod_tc(&av, b, &rv, t_stream));
    bump(result);
}
template <typename T>
void mod_ct(int64_t a, const TensorPtr<T>& b, TensorPtr<T>& result) {
    require(b, "b"), require(result, "result");
    hx_view bv = HX_V(b), rv = HX_V(result);
    check(hx_mod_ct(a, &bv, &rv, t_stream));
    bump(result);
}

template <typename T>
void axis_modsum(const TensorPtr<T>& a, const TensorPtr<T>& p, TensorPtr<T>& result, int64_t axis) {
    require(a, "a"), require(p, "p"), require(result, "result");
    hx_view av = HX_V(a), pv = HX_V(p), rv = HX_V(result);
    check(hx_axis_modsum(&av, &pv, &rv, axis, t_stream));
    bump(result);
}

template <typename T>
void g_decomposition(const TensorPtr<T>& a, TensorPtr<T>& result, size_t power, size_t base_bits) {
    require(a, "a"), require(result, "result");
    hx_view av = HX_V(a), rv = HX_V(result);
    uint64_t tag;
    {
        std::lock_guard<std::mutex> lk(g_warn.mu);
        tag = ++g_warn.next_tag;
        g_warn.calls[tag] = GdCall{a->dims, power, base_bits};
        while (g_warn.calls.size() > PendingWarn::kKeep) g_warn.calls.erase(g_warn.calls.begin());
    }
    check(hx_g_decomposition_tagged(&av, &rv, power, base_bits, tag, t_stream));
    bump(result);
}

// ntt / intt: the reference validates a, result, p, perm and twiddles in this
// order (example_impl/ntt_impl.cpp:24-42); log2p_list and mu_list are unused
template <typename T>
static void ntt_common(int inverse, const TensorPtr<T>& a, const TensorPtr<T>& p, const TensorPtr<T>& perm,
                       const TensorPtr<T>& tw, const TensorPtr<T>& minv, TensorPtr<T>& result) {
    require(a, "a"), require(result, "result"), require(p, "p"), require(perm, "perm"), require(tw, "twiddles");
    if (a->dims.size() != 4) throw std::invalid_argument("Input tensor 'a' must have shape [l, m, r, k].");
    if (result->dims != a->dims) throw std::invalid_argument("Output tensor must have the same shape as input tensor.");
    if (p->dims.size() != 1 || p->dims[0] != a->dims[3]) throw std::invalid_argument("Tensor 'p' must have shape [k].");
    if (perm->dims.size() != 1 || perm->dims[0] != a->dims[1])
        throw std::invalid_argument("Tensor 'perm' must have shape [m].");
    if (tw->dims.size() != 2 || tw->dims[0] != a->dims[3] || tw->dims[1] != a->dims[1])
        throw std::invalid_argument("Tensor 'twiddles' must have shape [k, m].");
    if (inverse) require(minv, "m_inv");
    auto plan = get_plan<T>(inverse, p, perm, tw, inverse ? minv : TensorPtr<T>());
    pin_if_capturing(plan);
    hx_view av = HX_V(a), rv = HX_V(result);
    check(hx_ntt_plan_exec(plan->plan, &av, &rv, t_stream));
    bump(result);
}

template <typename T>
void ntt(const TensorPtr<T>& a, const TensorPtr<T>& p, const TensorPtr<T>& perm, const TensorPtr<T>& twiddles,
         const TensorPtr<T>& log2p_list, const TensorPtr<T>& mu_list, TensorPtr<T>& result) {
    (void)log2p_list;
    (void)mu_list;
    ntt_common<T>(0, a, p, perm, twiddles, TensorPtr<T>(), result);
}

template <typename T>
void intt(const TensorPtr<T>& a, const TensorPtr<T>& p, const TensorPtr<T>& perm, const TensorPtr<T>& inv_twiddles,
          const TensorPtr<T>& m_inv, const TensorPtr<T>& log2p_list, const TensorPtr<T>& mu_list,
          TensorPtr<T>& result) {
    (void)log2p_list;
    (void)mu_list;
    ntt_common<T>(1, a, p, perm, inv_twiddles, m_inv, result);
}

// permute: the first use of a perms table (buffer, version, view, axes) is
// validated synchronously so that out-of-range entries throw from this call
template <typename T>
void permute(const TensorPtr<T>& a, const TensorPtr<T>& perms, TensorPtr<T>& result, int64_t elementwise_axis,
             int64_t perm_axis) {
    require(a, "a"), require(perms, "perms"), require(result, "result");
    hx_view av = HX_V(a), pv = HX_V(perms), rv = HX_V(result);
    const TensorKey key = key_of(perms);  // entries are checked against perms.shape[1]
    bool known = false, rows_same = false;
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto it = g_valid_perms.find(key);
        if (!a->dims.empty() && it != g_valid_perms.end()) {
            known = true;
            rows_same = it->second;
        }
    }
    if (known) {
        const uint32_t flags = HX_PERMS_TRUSTED | (rows_same ? HX_PERMS_ROWS_SAME : 0u);
        check(hx_permute_ex(&av, &pv, &rv, elementwise_axis, perm_axis, flags, t_stream));
    } else {
        check(hx_permute(&av, &pv, &rv, elementwise_axis, perm_axis, t_stream));
        if (!a->dims.empty()) {
            int same = 0;
            check(hx_perms_rows_same(&pv, &same, t_stream));
            std::lock_guard<std::mutex> lk(g_cache_mu);
            if (g_valid_perms.size() > 1024) g_valid_perms.clear();
            g_valid_perms[key] = same != 0;
        }
    }
    bump(result);
}

// ------------------------------------------------------------------ extensions
// Transcript operations named by the reference runtime without a reference
// definition (python_execution/lattica_heal_runtime/device_interface.py).

template <typename T>
void modneg_tt(const TensorPtr<T>& a, const TensorPtr<T>& p, TensorPtr<T>& result) {
    require(a, "a");
    require(p, "p");
    require(result, "result");
    hx_view av = HX_V(a), pv = HX_V(p), rv = HX_V(result);
    check(hx_modneg_tt(&av, &pv, &rv, t_stream));
    bump(result);
}

template <typename T>
void modneg_tc(const TensorPtr<T>& a, T p_scalar, TensorPtr<T>& result) {
    require(a, "a");
    require(result, "result");
    hx_view av = HX_V(a), rv = HX_V(result);
    check(hx_modneg_tc(&av, (int64_t)p_scalar, &rv, t_stream));
    bump(result);
}

template <typename T>
void modmul_axis_sum(const TensorPtr<T>& a, const TensorPtr<T>& b, const TensorPtr<T>& p, TensorPtr<T>& result,
                     int64_t axis) {
    require(a, "a");
    require(b, "b");
    require(p, "p");
    require(result, "result");
    hx_view av = HX_V(a), bv = HX_V(b), pv = HX_V(p), rv = HX_V(result);
    check(hx_modmul_axis_sum(&av, &bv, &pv, &rv, axis, t_stream));
    bump(result);
}

template <typename T>
void abs_values(const TensorPtr<T>& a, TensorPtr<T>& result) {
    require(a, "a");
    require(result, "result");
    hx_view av = HX_V(a), rv = HX_V(result);
    check(hx_abs(&av, &rv, t_stream));
    bump(result);
}

template <typename T>
void set_const_val(const TensorPtr<T>& a, T value) {
    require(a, "a");
    hx_view av = HX_V(a);
    check(hx_set_const(&av, std::is_floating_point<T>::value ? 0 : (int64_t)value, (double)value, t_stream));
    bump(a);
}

static int64_t norm_axis(int64_t axis, int64_t nd, const char* what) {
    if (axis < -nd || axis >= nd) throw std::invalid_argument(std::string(what) + " out of range");
    return axis < 0 ? axis + nd : axis;
}

template <typename T>
TensorPtr<T> moveaxis(const TensorPtr<T>& a, int64_t source, int64_t destination) {
    require(a, "a");
    const int64_t nd = (int64_t)a->dims.size();
    const int64_t s = norm_axis(source, nd, "source axis"), d = norm_axis(destination, nd, "destination axis");
    std::vector<int64_t> order;
    for (int64_t i = 0; i < nd; ++i)
        if (i != s) order.push_back(i);
    order.insert(order.begin() + d, s);
    DeviceTensor<T> src = *a;
    for (int64_t i = 0; i < nd; ++i) {
        src.dims[i] = a->dims[order[i]];
        src.strides[i] = a->strides[order[i]];
    }
    auto out = allocate_on_hardware<T>(src.dims);
    hx_view sv = view_of(src), ov = HX_V(out);
    check(hx_strided_copy(&sv, &ov, t_stream));
    return out;
}

template <typename T>
TensorPtr<T> get_slice(const TensorPtr<T>& a, const std::vector<SliceSpec>& spec) {
    require(a, "a");
    if (spec.size() > a->dims.size()) throw std::out_of_range("too many indices for tensor");
    auto v = std::make_shared<DeviceTensor<T>>(*a);
    v->dims.clear();
    v->strides.clear();
    for (size_t i = 0; i < a->dims.size(); ++i) {
        const int64_t dim = a->dims[i], st = a->strides[i];
        if (i >= spec.size()) {
            v->dims.push_back(dim);
            v->strides.push_back(st);
            continue;
        }
        const SliceSpec& sp = spec[i];
        if (sp.index) {
            if (sp.start < 0 || sp.start >= dim) throw std::out_of_range("index out of range");
            v->offset += sp.start * st;
            continue;
        }
        if (sp.step <= 0) throw std::invalid_argument("slice step must be positive");
        if (sp.start < 0 || sp.stop > dim || sp.start > sp.stop) throw std::out_of_range("slice out of range");
        const int64_t n = (sp.stop - sp.start + sp.step - 1) / sp.step;
        v->offset += (n ? sp.start * st : 0);
        v->dims.push_back(n);
        v->strides.push_back(st * sp.step);
    }
    return v;
}

template <typename T>
TensorPtr<T> pad_single_axis(const TensorPtr<T>& a, int64_t axis, int64_t before, int64_t after) {
    require(a, "a");
    const int64_t nd = (int64_t)a->dims.size();
    const int64_t ax = norm_axis(axis, nd, "axis");
    if (before < 0 || after < 0) throw std::invalid_argument("padding must be non-negative");
    std::vector<int64_t> dims = a->dims;
    dims[ax] += before + after;
    auto out = allocate_on_hardware<T>(dims);
    DeviceTensor<T> inner = *out;  // the sub-view that receives a
    inner.dims = a->dims;
    inner.offset = before * out->strides[ax];
    hx_view sv = HX_V(a), dv = view_of(inner);
    check(hx_strided_copy(&sv, &dv, t_stream));
    return out;
}

template <typename T>
TensorPtr<T> alias_view(const TensorPtr<T>& a, const std::vector<int64_t>& dims, const std::vector<int64_t>& strides,
                        int64_t offset) {
    require(a, "a");
    if (dims.size() != strides.size()) throw std::invalid_argument("dims and strides differ in length");
    int64_t lo = offset, hi = offset;
    for (size_t i = 0; i < dims.size(); ++i) {
        if (dims[i] < 0) throw std::invalid_argument("dimensions must be non-negative");
        if (dims[i] == 0) {
            lo = hi = offset;
            break;
        }
        if (strides[i] >= 0) hi += (dims[i] - 1) * strides[i];
        else lo += (dims[i] - 1) * strides[i];
    }
    const int64_t cap = a->buffer ? (int64_t)(a->buffer->bytes / sizeof(T)) : 0;
    if (lo < 0 || (hi >= cap && cap > 0) || (cap == 0 && hi > 0)) throw std::out_of_range("view outside its buffer");
    auto v = std::make_shared<DeviceTensor<T>>();
    v->dims = dims;
    v->strides = strides;
    v->buffer = a->buffer;
    v->offset = offset;
    return v;
}

template <typename T>
void copy_from_host(const TensorPtr<T>& a, const torch::Tensor& tensor) {
    require(a, "a");
    if (tensor.scalar_type() != torch::CppTypeToScalarType<T>())
        throw std::runtime_error("Tensor dtype does not match template parameter T.");
    if (std::vector<int64_t>(tensor.sizes().begin(), tensor.sizes().end()) != a->dims ||
        std::vector<int64_t>(tensor.strides().begin(), tensor.strides().end()) != a->strides)
        throw std::invalid_argument("host tensor must have the device tensor's dims and strides");
    const int64_t span = span_elems(a->dims, a->strides);
    if (span) {
        if (tensor.is_cuda())
            check(hx_memcpy_d2d(a->data_ptr(), tensor.data_ptr(), (size_t)span * sizeof(T), t_stream));
        else
            check(hx_memcpy_h2d(a->data_ptr(), tensor.data_ptr(), (size_t)span * sizeof(T), t_stream));
    }
    bump(a);
}

// ------------------------------------------------------------------ instantiations
#define HX_INST_MEM(T)                                                                                          \
    template std::shared_ptr<DeviceTensor<T>> allocate_on_hardware<T>(const std::vector<int64_t>&);            \
    template std::shared_ptr<DeviceTensor<T>> host_to_device<T>(const torch::Tensor&);                          \
    template torch::Tensor device_to_host<T>(const std::shared_ptr<DeviceTensor<T>>&);                          \
    template std::shared_ptr<DeviceTensor<T>> expand<T>(const std::shared_ptr<DeviceTensor<T>>&, int64_t, int64_t); \
    template std::shared_ptr<DeviceTensor<T>> squeeze<T>(const std::shared_ptr<DeviceTensor<T>>&, int64_t);   \
    template std::shared_ptr<DeviceTensor<T>> unsqueeze<T>(const std::shared_ptr<DeviceTensor<T>>&, int64_t); \
    template std::shared_ptr<DeviceTensor<T>> make_contiguous<T>(const std::shared_ptr<DeviceTensor<T>>&);    \
    template void set_const_val<T>(const TensorPtr<T>&, T);                                                   \
    template TensorPtr<T> moveaxis<T>(const TensorPtr<T>&, int64_t, int64_t);                                 \
    template TensorPtr<T> get_slice<T>(const TensorPtr<T>&, const std::vector<SliceSpec>&);                  \
    template TensorPtr<T> pad_single_axis<T>(const TensorPtr<T>&, int64_t, int64_t, int64_t);                \
    template TensorPtr<T> alias_view<T>(const TensorPtr<T>&, const std::vector<int64_t>&,                     \
                                        const std::vector<int64_t>&, int64_t);                                \
    template void copy_from_host<T>(const TensorPtr<T>&, const torch::Tensor&);

#define HX_INST_OPS(T)                                                                                     \
    template void modmul_ttt<T>(const TensorPtr<T>&, const TensorPtr<T>&, const TensorPtr<T>&, TensorPtr<T>&); \
    template void modmul_ttc<T>(const TensorPtr<T>&, const TensorPtr<T>&, T, TensorPtr<T>&);               \
    template void modmul_tct<T>(const TensorPtr<T>&, T, const TensorPtr<T>&, TensorPtr<T>&);               \
    template void modmul_tcc<T>(const TensorPtr<T>&, T, T, TensorPtr<T>&);                                 \
    template void modsum_ttt<T>(const TensorPtr<T>&, const TensorPtr<T>&, const TensorPtr<T>&, TensorPtr<T>&); \
    template void modsum_ttc<T>(const TensorPtr<T>&, const TensorPtr<T>&, T, TensorPtr<T>&);               \
    template void modsum_tct<T>(const TensorPtr<T>&, T, const TensorPtr<T>&, TensorPtr<T>&);               \
    template void modsum_tcc<T>(const TensorPtr<T>&, T, T, TensorPtr<T>&);                                 \
    template void mod_tt<T>(const TensorPtr<T>&, const TensorPtr<T>&, TensorPtr<T>&);                      \
    template void mod_tc<T>(const TensorPtr<T>&, int64_t, TensorPtr<T>&);                                  \
    template void mod_ct<T>(int64_t, const TensorPtr<T>&, TensorPtr<T>&);                                  \
    template void axis_modsum<T>(const TensorPtr<T>&, const TensorPtr<T>&, TensorPtr<T>&, int64_t);        \
    template void g_decomposition<T>(const TensorPtr<T>&, TensorPtr<T>&, size_t, size_t);                  \
    template void ntt<T>(const TensorPtr<T>&, const TensorPtr<T>&, const TensorPtr<T>&, const TensorPtr<T>&, \
                         const TensorPtr<T>&, const TensorPtr<T>&, TensorPtr<T>&);                         \
    template void intt<T>(const TensorPtr<T>&, const TensorPtr<T>&, const TensorPtr<T>&, const TensorPtr<T>&, \
                          const TensorPtr<T>&, const TensorPtr<T>&, const TensorPtr<T>&, TensorPtr<T>&);   \
    template void permute<T>(const TensorPtr<T>&, const TensorPtr<T>&, TensorPtr<T>&, int64_t, int64_t);    \
    template void modneg_tt<T>(const TensorPtr<T>&, const TensorPtr<T>&, TensorPtr<T>&);                   \
    template void modneg_tc<T>(const TensorPtr<T>&, T, TensorPtr<T>&);                                     \
    template void modmul_axis_sum<T>(const TensorPtr<T>&, const TensorPtr<T>&, const TensorPtr<T>&, TensorPtr<T>&, \
                                     int64_t);                                                            \
    template void abs_values<T>(const TensorPtr<T>&, TensorPtr<T>&);

HX_INST_MEM(int32_t)
HX_INST_MEM(int64_t)
HX_INST_MEM(double)
HX_INST_OPS(int32_t)
HX_INST_OPS(int64_t)

}  // namespace lattica_hw_api
