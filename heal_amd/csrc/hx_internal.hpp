// hx_internal.hpp — shared host/device plumbing of the HEAL MI355X backend:
// error reporting at the C-ABI boundary, view validation, iteration-shape
// coalescing for strided kernels, fast 32-bit division by invariants, and the
// per-device status block that kernels use to report data-dependent errors.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/heal_amd.h"
#include "hx_arith.hpp"

namespace hx {

// ------------------------------------------------------------------ errors
struct Error {
    int code;
    std::string msg;
};

[[noreturn]] void raise(int code, const std::string& msg);
int set_error(int code, const std::string& msg);
void check_hip(hipError_t e, const char* what);

#define HX_API_BEGIN try {
#define HX_API_END                                                        \
    }                                                                     \
    catch (const hx::Error& e) {                                          \
        return hx::set_error(e.code, e.msg);                              \
    }                                                                     \
    catch (const std::bad_alloc&) {                                       \
        return hx::set_error(HX_ERR_BAD_ALLOC, "std::bad_alloc");         \
    }                                                                     \
    catch (const std::exception& e) {                                     \
        return hx::set_error(HX_ERR_RUNTIME, e.what());                   \
    }                                                                     \
    return HX_OK;

// ------------------------------------------------------------------- views
size_t dtype_size(int dtype);
void check_view(const hx_view* v, const char* name);  // throws on malformed descriptors
int64_t view_numel(const hx_view* v);
bool dims_equal(const hx_view* a, const hx_view* b);
// [lowest, highest] byte range touched by a view (empty views give lo > hi).
void view_byte_range(const hx_view* v, uintptr_t* lo, uintptr_t* hi);
bool views_overlap(const hx_view* a, const hx_view* b);

// ------------------------------------------- fast division by an invariant
// Granlund & Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994 (round-up variant): for n < 2^31,
// n / d = (umulhi(n, mul) + n) >> shift.
#if defined(__HIPCC__)
// The reference's modmul (op 0) / modsum (op 1) of one element pair,
// (T_DP(a) op T_DP(b)) % p with |p| described by mi (example_impl/modop_impl.cpp:
// 123-126): operands that are residues (0 <= x < |p|) take one double-width
// product and reduction, or one conditional subtraction; any other pair takes
// the truncating signed semantics.  A modulus of 0 gives 0.
__device__ __forceinline__ uint64_t modop_fast(int op, const ModInfo& mi, int64_t av, int64_t bv) {
    if (mi.M == 0) return 0;
    if ((uint64_t)av < mi.M && (uint64_t)bv < mi.M) {
        if (op == 0) return mulmod(mi, (uint64_t)av, (uint64_t)bv);
        const uint64_t s2 = (uint64_t)av + (uint64_t)bv;
        return s2 >= mi.M ? s2 - mi.M : s2;
    }
    return (uint64_t)(op == 0 ? cpp_mulmod(mi, av, bv) : cpp_addmod(mi, av, bv));
}
#endif

struct FastDiv {
    uint32_t d;
    uint32_t mul;
    uint32_t shift;
    uint32_t pad;
};

inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f;
    f.d = d;
    uint32_t l = 0;
    while (l < 32 && (1ull << l) < d) ++l;
    f.shift = l;
    f.mul = (uint32_t)((((1ull << l) - d) << 32) / d + 1);
    f.pad = 0;
    return f;
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv& f) {
    return (__umulhi(n, f.mul) + n) >> f.shift;
}

// ------------------------------------------------------- iteration shapes
// A strided iteration space over up to HX_MAX_DIMS dims with up to 4 operands.
// Size-1 dims are dropped and adjacent dims that are contiguous for every
// operand are merged, so kernels decode as few coordinates as possible.
constexpr int kMaxOps = 4;

struct IterSpace {
    int32_t ndim;
    int32_t use32;  // total < 2^31: 32-bit fast division
    int64_t total;
    int64_t dims[HX_MAX_DIMS];
    FastDiv fd[HX_MAX_DIMS];
    int64_t strides[kMaxOps][HX_MAX_DIMS];
};

// dims: iteration dims (ndim entries); strides[o][i]: stride of operand o along
// dim i (0 for broadcast).  keep_last: never merge the last dim away nor move it
// (used when a kernel indexes a table by the last coordinate).  The dims are
// walked in the memory order of operand order_op (its largest stride
// outermost; ties keep their order), so that a result laid out in another
// order than its logical one is still written, and read operands with the
// same layout are still read, in whole lines.
IterSpace make_iter(int ndim, const int64_t* dims, int nops, const int64_t (*strides)[HX_MAX_DIMS], bool keep_last,
                    int order_op = 0);

// Decode a linear index into per-operand element offsets; also returns the
// coordinate along the last iteration dim.
template <int NOPS>
__device__ __forceinline__ void iter_offsets(const IterSpace& it, int64_t idx, int64_t* off, int64_t* last) {
#pragma unroll
    for (int o = 0; o < NOPS; ++o) off[o] = 0;
    if (it.use32) {
        uint32_t n = (uint32_t)idx;
        for (int d = it.ndim - 1; d >= 0; --d) {
            uint32_t q = fdiv(n, it.fd[d]);
            uint32_t c = n - q * it.fd[d].d;
            if (d == it.ndim - 1) *last = c;
#pragma unroll
            for (int o = 0; o < NOPS; ++o) off[o] += (int64_t)c * it.strides[o][d];
            n = q;
        }
    } else {
        for (int d = it.ndim - 1; d >= 0; --d) {
            int64_t c = idx % it.dims[d];
            if (d == it.ndim - 1) *last = c;
#pragma unroll
            for (int o = 0; o < NOPS; ++o) off[o] += c * it.strides[o][d];
            idx /= it.dims[d];
        }
    }
}

// ------------------------------------------------------- device status
constexpr int kGdLog = 64;  // = HX_GDECOMP_LOG_MAX
struct DevStatus {
    unsigned long long gdecomp_overflow;
    unsigned long long index_out_of_range;
    unsigned long long gd_log_count;  // entries claimed in gd_log (may pass kGdLog; only kGdLog are kept)
    unsigned long long spare[5];
    unsigned long long gd_log[2 * kGdLog];  // (tag, row-major index in a) of overflowing elements
};
DevStatus* device_status();  // current device's block (allocated on first use)

// Records one g_decomposition element whose value was not exhausted; a
// work-item calls it at most once, and not at all once the log is full.
__device__ __forceinline__ void gd_log_element(DevStatus* st, unsigned long long tag, unsigned long long index) {
    if (*(volatile unsigned long long*)&st->gd_log_count >= (unsigned long long)kGdLog) return;
    const unsigned long long k = atomicAdd(&st->gd_log_count, 1ull);
    if (k < (unsigned long long)kGdLog) {
        st->gd_log[2 * k] = tag;
        st->gd_log[2 * k + 1] = index;
    }
}

// grid size for grid-stride kernels
inline unsigned grid_for(int64_t total, int block, int64_t cap = 4096) {
    int64_t g = (total + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

inline hipStream_t as_stream(hx_stream s) { return (hipStream_t)s; }

void fill_zero(void* ptr, size_t bytes, hipStream_t st);     // fill kernel (graph-safe)
void fill_u32(uint32_t* ptr, int64_t n, uint32_t value, hipStream_t st);

// ------------------------------------------------ streams, capture, release
// True while work enqueued on st is being captured into a graph.
bool stream_capturing(hipStream_t st);

// Work that needs a host synchronisation (plan preparation, table checks).
// It runs on the caller's stream, or, when that stream is being captured, on
// the current device's side stream with the thread's capture mode relaxed, so
// that the capture stays valid.  Tables read on the side stream must be
// resident before the capture began.
class SyncScope {
  public:
    explicit SyncScope(hipStream_t st);
    ~SyncScope();
    SyncScope(const SyncScope&) = delete;
    SyncScope& operator=(const SyncScope&) = delete;
    hipStream_t stream() const { return st_; }
    void sync();

  private:
    hipStream_t st_;
    bool escaped_;
    int mode_;
};

// Frees pool memory (hipMallocAsync) once every piece of work marked by
// `events` has run: the device's reclaim stream waits on the events and frees
// in its own order, so the host never waits.  Destroys the events.
void release_after(int device, std::vector<hipEvent_t>& events, const std::vector<void*>& ptrs);

// Workspaces of one plan.  Eager execution gets one workspace per stream,
// reused and grown in that stream's order, so plans can be shared by threads
// and streams.  An execution that is being captured into a graph gets a
// workspace of its own that lives as long as the plan (the graph keeps using
// it).  A workspace has a `flags` part, zero when created, that its users
// leave zero after every execution, and a `data` part.
struct Workspaces {
    struct Ws {
        void* flags = nullptr;
        size_t flag_bytes = 0;
        void* data = nullptr;
        size_t data_bytes = 0;
    };
    std::mutex mu;  // held from acquire() to note()
    struct Eager {
        hipStream_t s;
        Ws ws;
        hipEvent_t ev;
    };
    std::vector<Eager> eager;
    std::vector<Ws> captured;
    int device = 0;
    // A workspace for work about to be enqueued on st (callers hold mu).
    Ws acquire(hipStream_t st, size_t flag_bytes, size_t data_bytes);
    // After that work is enqueued: marks its completion on st (callers hold mu).
    void note(hipStream_t st);
    // All memory events of eager use, plus the workspaces themselves, for release_after.
    void release(std::vector<hipEvent_t>& events, std::vector<void*>& ptrs);
};

}  // namespace hx
