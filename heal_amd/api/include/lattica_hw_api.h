// lattica_hw_api.h — the HEAL hardware API on MI355X, C++17 side.
//
// Declarations keep the signatures of the reference headers (Lattica-ai/heal
// include/*.h); the implementations call the C ABI of include/heal_amd.h.
// All ops enqueue work on the library stream (set_stream) and return; device
// results become host-visible through device_to_host / print, which
// synchronise.  Errors are thrown as the reference throws them:
// std::invalid_argument, std::out_of_range, std::runtime_error, std::bad_alloc.
#ifndef HEAL_AMD_LATTICA_HW_API_H
#define HEAL_AMD_LATTICA_HW_API_H

#include "device_memory.h"

namespace lattica_hw_api {

template <typename T>
using TensorPtr = std::shared_ptr<DeviceTensor<T>>;

// ---- views (reference include/memory_virtual_ops.h:39-86, include/contiguous.h:19-20)
template <typename T>
std::shared_ptr<DeviceTensor<T>> expand(const std::shared_ptr<DeviceTensor<T>>& tensor, int64_t axis, int64_t repeats);
template <typename T>
std::shared_ptr<DeviceTensor<T>> squeeze(const std::shared_ptr<DeviceTensor<T>>& tensor, int64_t axis);
template <typename T>
std::shared_ptr<DeviceTensor<T>> unsqueeze(const std::shared_ptr<DeviceTensor<T>>& tensor, int64_t axis);
template <typename T>
std::shared_ptr<DeviceTensor<T>> make_contiguous(const std::shared_ptr<DeviceTensor<T>>& tensor);

// ---- modular elementwise (reference include/modop.h:47-134)
template <typename T>
void modmul_ttt(const std::shared_ptr<DeviceTensor<T>>& a, const std::shared_ptr<DeviceTensor<T>>& b,
                const std::shared_ptr<DeviceTensor<T>>& p, std::shared_ptr<DeviceTensor<T>>& result);
template <typename T>
void modmul_ttc(const std::shared_ptr<DeviceTensor<T>>& a, const std::shared_ptr<DeviceTensor<T>>& b, T p_scalar,
                std::shared_ptr<DeviceTensor<T>>& result);
template <typename T>
void modmul_tct(const std::shared_ptr<DeviceTensor<T>>& a, T b_scalar, const std::shared_ptr<DeviceTensor<T>>& p,
                std::shared_ptr<DeviceTensor<T>>& result);
template <typename T>
void modmul_tcc(const std::shared_ptr<DeviceTensor<T>>& a, T b_scalar, T p_scalar,
                std::shared_ptr<DeviceTensor<T>>& result);
template <typename T>
void modsum_ttt(const std::shared_ptr<DeviceTensor<T>>& a, const std::shared_ptr<DeviceTensor<T>>& b,
                const std::shared_ptr<DeviceTensor<T>>& p, std::shared_ptr<DeviceTensor<T>>& result);
template <typename T>
void modsum_ttc(const std::shared_ptr<DeviceTensor<T>>& a, const std::shared_ptr<DeviceTensor<T>>& b, T p_scalar,
                std::shared_ptr<DeviceTensor<T>>& result);
template <typename T>
void modsum_tct(const std::shared_ptr<DeviceTensor<T>>& a, T b_scalar, const std::shared_ptr<DeviceTensor<T>>& p,
                std::shared_ptr<DeviceTensor<T>>& result);
template <typename T>
void modsum_tcc(const std::shared_ptr<DeviceTensor<T>>& a, T b_scalar, T p_scalar,
                std::shared_ptr<DeviceTensor<T>>& result);
template <typename T>
void mod_tt(const std::shared_ptr<DeviceTensor<T>>& a, const std::shared_ptr<DeviceTensor<T>>& b,
            std::shared_ptr<DeviceTensor<T>>& result);
template <typename T>
void mod_tc(const std::shared_ptr<DeviceTensor<T>>& a, int64_t b_scalar, std::shared_ptr<DeviceTensor<T>>& result);
template <typename T>
void mod_ct(int64_t a_scalar, const std::shared_ptr<DeviceTensor<T>>& b, std::shared_ptr<DeviceTensor<T>>& result);

// ---- reduction over an axis (reference include/axis_modsum.h:24-30)
template <typename T>
void axis_modsum(const std::shared_ptr<DeviceTensor<T>>& a, const std::shared_ptr<DeviceTensor<T>>& p,
                 std::shared_ptr<DeviceTensor<T>>& result, int64_t axis);

// ---- gadget decomposition (reference include/g_decomposition.h:24-30)
template <typename T>
void g_decomposition(const std::shared_ptr<DeviceTensor<T>>& a, std::shared_ptr<DeviceTensor<T>>& result,
                     size_t power, size_t base_bits);

// ---- NTT / INTT (reference include/ntt.h:31-52)
template <typename T>
void ntt(const std::shared_ptr<DeviceTensor<T>>& a, const std::shared_ptr<DeviceTensor<T>>& p,
         const std::shared_ptr<DeviceTensor<T>>& perm, const std::shared_ptr<DeviceTensor<T>>& twiddles,
         const std::shared_ptr<DeviceTensor<T>>& log2p_list, const std::shared_ptr<DeviceTensor<T>>& mu_list,
         std::shared_ptr<DeviceTensor<T>>& result);
template <typename T>
void intt(const std::shared_ptr<DeviceTensor<T>>& a, const std::shared_ptr<DeviceTensor<T>>& p,
          const std::shared_ptr<DeviceTensor<T>>& perm, const std::shared_ptr<DeviceTensor<T>>& inv_twiddles,
          const std::shared_ptr<DeviceTensor<T>>& m_inv, const std::shared_ptr<DeviceTensor<T>>& log2p_list,
          const std::shared_ptr<DeviceTensor<T>>& mu_list, std::shared_ptr<DeviceTensor<T>>& result);

// ---- permutations (reference include/permute.h:33-40)
template <typename T>
void permute(const std::shared_ptr<DeviceTensor<T>>& a, const std::shared_ptr<DeviceTensor<T>>& perms,
             std::shared_ptr<DeviceTensor<T>>& result, int64_t elementwise_axis, int64_t perm_axis);

// ---- transcript operations that the reference's runtime names but no reference
// file defines (python_execution/lattica_heal_runtime/device_interface.py:47-48,
// 75-107).  Semantics are this library's own (stated per function).

// result = the residue of -a modulo |p| in [0, |p|) (0 where p == 0); p broadcasts as in modmul
template <typename T>
void modneg_tt(const std::shared_ptr<DeviceTensor<T>>& a, const std::shared_ptr<DeviceTensor<T>>& p,
               std::shared_ptr<DeviceTensor<T>>& result);
template <typename T>
void modneg_tc(const std::shared_ptr<DeviceTensor<T>>& a, T p_scalar, std::shared_ptr<DeviceTensor<T>>& result);
// result = axis_modsum(modmul_ttt(a, b, p), p, axis) without the product tensor; a and b
// broadcast against each other (right-aligned) and axis counts in the broadcast shape
template <typename T>
void modmul_axis_sum(const std::shared_ptr<DeviceTensor<T>>& a, const std::shared_ptr<DeviceTensor<T>>& b,
                     const std::shared_ptr<DeviceTensor<T>>& p, std::shared_ptr<DeviceTensor<T>>& result,
                     int64_t axis);
// result = |a| (two's complement wrap at the minimum)
template <typename T>
void abs_values(const std::shared_ptr<DeviceTensor<T>>& a, std::shared_ptr<DeviceTensor<T>>& result);
// every element of the view := value (the buffer is written in place)
template <typename T>
void set_const_val(const std::shared_ptr<DeviceTensor<T>>& a, T value);
// numpy.moveaxis materialised: a new C-contiguous tensor
template <typename T>
std::shared_ptr<DeviceTensor<T>> moveaxis(const std::shared_ptr<DeviceTensor<T>>& a, int64_t source,
                                          int64_t destination);
// Basic indexing of the leading dims, numpy style: a view sharing a's buffer.
// A spec with index == true takes element `start` and drops the dim; otherwise
// it keeps start, start + step, ... below stop (0 <= start <= stop <= dim, step > 0).
struct SliceSpec {
    int64_t start = 0, stop = 0, step = 1;
    bool index = false;
};
template <typename T>
std::shared_ptr<DeviceTensor<T>> get_slice(const std::shared_ptr<DeviceTensor<T>>& a,
                                           const std::vector<SliceSpec>& spec);
// a new C-contiguous tensor with `before` zeros ahead of and `after` zeros behind
// the data along `axis`
template <typename T>
std::shared_ptr<DeviceTensor<T>> pad_single_axis(const std::shared_ptr<DeviceTensor<T>>& a, int64_t axis,
                                                 int64_t before, int64_t after);

// ---- views and transfers for programs replayed many times (no counterpart in
// the reference)
// A new tensor object over a's buffer with the given metadata (no copy); the
// addressed span must lie inside the buffer.
template <typename T>
std::shared_ptr<DeviceTensor<T>> alias_view(const std::shared_ptr<DeviceTensor<T>>& a, const std::vector<int64_t>& dims,
                                            const std::vector<int64_t>& strides, int64_t offset);
// Writes a host tensor of dtype T with exactly a's dims and strides into a's
// buffer (the span it addresses) and marks the buffer as written.
template <typename T>
void copy_from_host(const std::shared_ptr<DeviceTensor<T>>& a, const torch::Tensor& tensor);

// ---- backend controls (no counterpart in the reference)
// Stream that subsequent ops of the calling thread are enqueued on (a
// hipStream_t; nullptr = the legacy default stream).
void set_stream(void* stream);
void* get_stream();
// Waits for the stream and reports deferred data-dependent conditions
// (g_decomposition overflow warnings on stderr).
void synchronize();
// Number of cached NTT plans / validated permutation tables (for tests).
size_t plan_cache_size();

// ---- graph capture (no counterpart in the reference)
// A graph captured from ops keeps addressing the device data of the NTT plans
// those ops used, so the plans must outlive the graph.  capture_begin() starts
// a record, for the calling thread, of every plan used while its stream is
// being captured; capture_end() returns the record, which the caller keeps as
// long as the graph (dropping it releases the plans; the graph must have
// finished running).  Plans used by a capture outside such a record are kept
// until the process ends.  Plan creation and table checks that would
// synchronise run on a side stream during capture, so tables must be resident
// before the capture begins.
struct CapturePins;
void capture_begin();
std::shared_ptr<CapturePins> capture_end();
// Plans currently held by captures (records and process-lifetime pins; for tests).
size_t captured_plan_count();

}  // namespace lattica_hw_api

#endif  // HEAL_AMD_LATTICA_HW_API_H
