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

}  // namespace lattica_hw_api

#endif  // HEAL_AMD_LATTICA_HW_API_H
