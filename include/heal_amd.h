/*
 * heal_amd.h — C ABI of the MI355X (gfx950) HEAL backend.
 *
 * Every compute entry point of the HEAL hardware API (reference: the
 * include/ headers of Lattica-ai/heal) has one extern "C" function here.  Tensors cross the boundary
 * as plain view descriptors (device pointer, dtype, dims, strides in elements);
 * no C++ or torch type appears in any signature.  The C++17 templates of the HEAL
 * API (namespace lattica_hw_api) and the pybind11 module `lattica_hw` are built on
 * top of this layer.
 *
 * Conventions
 *  - All functions return an hx_status.  On failure hx_last_error() gives the
 *    message (thread-local).  HX_ERR_INVALID_ARGUMENT / HX_ERR_OUT_OF_RANGE /
 *    HX_ERR_RUNTIME / HX_ERR_BAD_ALLOC correspond to std::invalid_argument,
 *    std::out_of_range, std::runtime_error and std::bad_alloc of the reference.
 *  - Work is enqueued on `stream` (a hipStream_t; NULL = the legacy default
 *    stream) and is asynchronous unless the description says otherwise.
 *  - A view's `data` points at element [0,...,0]; strides may be 0 (broadcast).
 *  - Element types: HX_INT32, HX_INT64 for compute ops; HX_FLOAT64 only for the
 *    memory/view ops, as in the reference.
 *  - Results are exact integers and equal the reference's wherever the
 *    reference is defined: C++ integer semantics (truncating remainder,
 *    128-bit/64-bit intermediates), including the NTT/INTT of inputs outside
 *    [0, q), whose signed representatives the reference's network keeps.
 *    Where the reference overflows (its INTT's m^-1 product in T) the result
 *    is the exact one.
 *  - Plans (NTT, key-switch, matrix-core NTT) may be used from several threads
 *    and streams at once: each stream gets its own workspace.  Destroying a
 *    plan never waits on the host; its memory is released after the work
 *    already queued with it.  A graph captured with a plan must be destroyed
 *    (and have finished) before the plan.
 */
#ifndef HEAL_AMD_H
#define HEAL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HX_MAX_DIMS 8
#define HX_ABI_VERSION 1

typedef enum hx_dtype { HX_INT32 = 0, HX_INT64 = 1, HX_FLOAT64 = 2 } hx_dtype;

typedef enum hx_status {
    HX_OK = 0,
    HX_ERR_INVALID_ARGUMENT = 1,
    HX_ERR_OUT_OF_RANGE = 2,
    HX_ERR_RUNTIME = 3,
    HX_ERR_BAD_ALLOC = 4
} hx_status;

typedef struct hx_view {
    void*   data;                  /* device address of element [0,...,0] */
    int32_t dtype;                 /* hx_dtype */
    int32_t ndim;                  /* 0..HX_MAX_DIMS */
    int64_t dims[HX_MAX_DIMS];
    int64_t strides[HX_MAX_DIMS];  /* in elements */
} hx_view;

typedef void* hx_stream;           /* hipStream_t */

/* Sticky per-device counters written by kernels for data-dependent conditions. */
typedef struct hx_device_status {
    uint64_t gdecomp_overflow;     /* g_decomposition elements whose value was not exhausted */
    uint64_t index_out_of_range;   /* permute / NTT perm entries outside [0, m) */
} hx_device_status;

/* ------------------------------------------------------------------ library */
int         hx_abi_version(void);
const char* hx_last_error(void);
/* Takes the current device's status counters in `stream`'s order (atomically
 * read and cleared after the stream's earlier work; counts of work still
 * running on other streams remain for a later poll), copies them into *out
 * (may be NULL) and synchronises `stream` only. */
int hx_status_poll(hx_stream stream, hx_device_status* out);

/* The elements behind gdecomp_overflow, for the reference's per-element
 * warning "Warning: value at <coords> exceeds capacity with base_bits=B and
 * power=P" (example_impl/g_decomposition_impl.cpp:58-65): up to
 * HX_GDECOMP_LOG_MAX entries (the call's tag from hx_g_decomposition_tagged,
 * 0 for hx_g_decomposition, and the element's row-major index in `a`), in no
 * particular order; further elements are counted in gdecomp_overflow only.
 * An entry written by work that runs on another stream while the poll takes
 * the log can be lost. */
#define HX_GDECOMP_LOG_MAX 64
typedef struct hx_gdecomp_log {
    uint64_t count;                      /* entries filled */
    uint64_t tag[HX_GDECOMP_LOG_MAX];
    uint64_t index[HX_GDECOMP_LOG_MAX];
} hx_gdecomp_log;

/* hx_status_poll that also takes the g_decomposition overflow log (log may be NULL). */
int hx_status_poll_log(hx_stream stream, hx_device_status* out, hx_gdecomp_log* log);

/* ---------------------------------------------------- memory (L0 of the API)
 * Replaces the storage side of DeviceTensor / allocate_on_hardware /
 * host_to_device / device_to_host
 * (reference include/device_memory.h:24-39, example_impl/device_memory_impl.cpp:168-202).
 * Allocation is stream-ordered from a caching pool. */
int hx_malloc(void** out, size_t bytes, hx_stream stream);
int hx_free(void* ptr, hx_stream stream);
int hx_memset_zero(void* ptr, size_t bytes, hx_stream stream);          /* fill kernel, graph-safe */
int hx_memcpy_h2d(void* dst, const void* src, size_t bytes, hx_stream stream); /* returns when src may be reused */
int hx_memcpy_d2h(void* dst, const void* src, size_t bytes, hx_stream stream); /* returns when dst is filled */
int hx_memcpy_d2d(void* dst, const void* src, size_t bytes, hx_stream stream);
int hx_stream_sync(hx_stream stream);
/* *capturing = 1 while work enqueued on `stream` is being captured into a graph.
 * No counterpart in the reference (its API has no streams). */
int hx_stream_capturing(hx_stream stream, int* capturing);

/* ------------------------------------------------------------ views (L1)
 * make_contiguous (reference include/contiguous.h:19-20, example_impl/contiguous_impl.cpp:9-52):
 * gathers the strided view `src` into `dst` (same dims, any strides, normally
 * C-contiguous).  Any dtype. */
int hx_strided_copy(const hx_view* src, const hx_view* dst, hx_stream stream);

/* ------------------------------------------------- modular elementwise (L2)
 * result = (a (*|+) b) % p with right-aligned broadcasting of a, b, p to result
 * (reference include/modop.h:47-111, example_impl/modop_impl.cpp:108-180).
 * Scalars are passed as int64 and converted to the element type. */
int hx_modmul_ttt(const hx_view* a, const hx_view* b, const hx_view* p, const hx_view* result, hx_stream stream);
int hx_modmul_ttc(const hx_view* a, const hx_view* b, int64_t p, const hx_view* result, hx_stream stream);
int hx_modmul_tct(const hx_view* a, int64_t b, const hx_view* p, const hx_view* result, hx_stream stream);
int hx_modmul_tcc(const hx_view* a, int64_t b, int64_t p, const hx_view* result, hx_stream stream);
int hx_modsum_ttt(const hx_view* a, const hx_view* b, const hx_view* p, const hx_view* result, hx_stream stream);
int hx_modsum_ttc(const hx_view* a, const hx_view* b, int64_t p, const hx_view* result, hx_stream stream);
int hx_modsum_tct(const hx_view* a, int64_t b, const hx_view* p, const hx_view* result, hx_stream stream);
int hx_modsum_tcc(const hx_view* a, int64_t b, int64_t p, const hx_view* result, hx_stream stream);

/* result = a % b, C++ truncating remainder, same shapes only
 * (reference include/modop.h:115-134, example_impl/modop_impl.cpp:183-235). */
int hx_mod_tt(const hx_view* a, const hx_view* b, const hx_view* result, hx_stream stream);
int hx_mod_tc(const hx_view* a, int64_t b, const hx_view* result, hx_stream stream);
int hx_mod_ct(int64_t a, const hx_view* b, const hx_view* result, hx_stream stream);

/* result = sum over `axis` of a, reduced mod p[k] after every term
 * (reference include/axis_modsum.h:24-30, example_impl/axis_modsum_impl.cpp:12-76). */
int hx_axis_modsum(const hx_view* a, const hx_view* p, const hx_view* result, int64_t axis, hx_stream stream);

/* Transcript operations that the reference's runtime names but no reference
 * file defines (python_execution/lattica_heal_runtime/device_interface.py:47-48,
 * 78-85: _modneg_tt/_modneg_tc, modmul_axis_sum, abs, set_const_val).
 * Their semantics here are this library's own:
 *   modneg:          result = the residue of -a modulo |p| in [0, |p|) (0 for p = 0);
 *                    p broadcasts as in modmul;
 *   modmul_axis_sum: result = axis_modsum(modmul_ttt(a, b, p), p, axis) with the
 *                    reference's step semantics of both ops, without the product
 *                    tensor; a and b broadcast against each other (right-aligned),
 *                    axis counts in the broadcast shape;
 *   abs:             result = |a| (two's complement wrap at the minimum);
 *   set_const:       every element of a := ivalue (integer dtypes) or fvalue (float64). */
int hx_modneg_tt(const hx_view* a, const hx_view* p, const hx_view* result, hx_stream stream);
int hx_modneg_tc(const hx_view* a, int64_t p, const hx_view* result, hx_stream stream);
int hx_modmul_axis_sum(const hx_view* a, const hx_view* b, const hx_view* p, const hx_view* result, int64_t axis,
                       hx_stream stream);
int hx_abs(const hx_view* a, const hx_view* result, hx_stream stream);

/* One step of the diagonal matrix-vector product on ciphertexts (the serving
 * loop's rotate-multiply-accumulate; no counterpart in the reference).  With
 * acc, ct and ks [B, 2, n, L], diags [B, M, M, L] (M a power of two), perm [n]
 * and p [L], all int64 and contiguous within an instance:
 *     rot_0 = modsum_ttt(ks[:, 0], permute(ct[:, 0], perm over rows)),  rot_1 = ks[:, 1]
 *     (rot = ct when ks is NULL; perm is then not read)
 *     acc[:, c] = modsum_ttt(acc[:, c], modmul_ttt(rot_c, pt)),
 *     pt[b, j, t] = diags[b, diag mod M, j mod M, t]
 * in one kernel, each step with the semantics of the op it names (diags holds
 * a matrix by diagonals: diags[b, r, i, t] = M_b[i, (i + r) mod M] mod q_t).
 * Entries of perm outside [0, n) read 0 and are counted as in permute
 * (reported by the next status poll).  acc must not overlap the other
 * operands; B * n * L < 2^31, L <= 512. */
int hx_rot_mac(const hx_view* acc, const hx_view* ks, const hx_view* ct, const hx_view* perm, const hx_view* diags,
               int64_t diag, const hx_view* p, hx_stream stream);
int hx_set_const(const hx_view* a, int64_t ivalue, double fvalue, hx_stream stream);

/* result[..., d] = d-th base-2^base_bits digit of a (unsigned digits)
 * (reference include/g_decomposition.h:24-30, example_impl/g_decomposition_impl.cpp:10-67).
 * Elements whose value is not exhausted after `power` digits increment the
 * gdecomp_overflow counter (the reference prints a warning). */
int hx_g_decomposition(const hx_view* a, const hx_view* result, uint64_t power, uint64_t base_bits, hx_stream stream);
/* hx_g_decomposition whose overflowing elements carry `tag` in the status log
 * (hx_status_poll_log), so that the caller can print the reference's
 * per-element warning with the call's shape, base_bits and power. */
int hx_g_decomposition_tagged(const hx_view* a, const hx_view* result, uint64_t power, uint64_t base_bits,
                              uint64_t tag, hx_stream stream);

/* Negacyclic NTT / INTT over axis 1 of [l, m, r, k] views
 * (reference include/ntt.h:31-52, example_impl/ntt_impl.cpp:69-170).
 * log2p_list and mu_list may be NULL and are not read (as in the reference).
 * These build a temporary plan (see hx_ntt_plan_*) and synchronise the stream
 * once to classify `perm`.
 * Columns whose reference result is not the canonical residue (negative
 * inputs, INTT inputs outside [0, q), forward inputs above 2^63 - 1 - q) are
 * recomputed by a kernel that replays the reference's signed network; tables
 * with entries outside [0, q), and moduli above 2^62 (int64) or, for the
 * int32 INTT, 2^30, send every column there. */
int hx_ntt(const hx_view* a, const hx_view* p, const hx_view* perm, const hx_view* twiddles,
           const hx_view* log2p_list, const hx_view* mu_list, const hx_view* result, hx_stream stream);
int hx_intt(const hx_view* a, const hx_view* p, const hx_view* perm, const hx_view* inv_twiddles,
            const hx_view* m_inv, const hx_view* log2p_list, const hx_view* mu_list,
            const hx_view* result, hx_stream stream);

/* result[c] = a[c with c[perm_axis] := perms[c[elementwise_axis], c[perm_axis]]]
 * (reference include/permute.h:33-40, example_impl/permute_impl.cpp:8-65).
 * Synchronises the stream to report out-of-range perms as HX_ERR_OUT_OF_RANGE. */
int hx_permute(const hx_view* a, const hx_view* perms, const hx_view* result,
               int64_t elementwise_axis, int64_t perm_axis, hx_stream stream);

/* ------------------------------------------------------ plans and batching
 * No counterpart in the reference: they let callers that replay the same
 * constants (twiddle tables, Galois maps) skip per-call preparation and host
 * synchronisation, which the HEAL API's per-call form cannot express.
 *
 * An NTT plan is an immutable device snapshot of (p, perm, twiddles[, m_inv])
 * with Shoup companions precomputed; executing it never reads the source
 * tensors again and never synchronises. */
typedef struct hx_ntt_plan hx_ntt_plan;
int hx_ntt_plan_create(hx_ntt_plan** out, int inverse, const hx_view* p, const hx_view* perm,
                       const hx_view* twiddles, const hx_view* m_inv, hx_stream stream);
int hx_ntt_plan_exec(const hx_ntt_plan* plan, const hx_view* a, const hx_view* result, hx_stream stream);
int hx_ntt_plan_destroy(hx_ntt_plan* plan);
/* hx_ntt_plan_exec with a caller-provided workspace (256-byte aligned, at
 * least the size hx_ntt_plan_workspace_bytes gives for the same views; its
 * contents on entry do not matter and it must not be shared by executions in
 * flight at the same time).  No counterpart in the reference. */
int hx_ntt_plan_workspace_bytes(const hx_ntt_plan* plan, const hx_view* a, const hx_view* result, size_t* bytes);
int hx_ntt_plan_exec_ws(const hx_ntt_plan* plan, const hx_view* a, const hx_view* result, void* workspace,
                        size_t workspace_bytes, hx_stream stream);
/* hx_ntt_plan_exec that allocates nothing: its workspace is a fixed device
 * array of HX_NTT_NOALLOC_SCRATCH_BYTES inside the library (one per GPU, held
 * until the process ends).  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Limit: the transform's workspace
 * (hx_ntt_plan_workspace_bytes: about 8 * m * columns bytes for two or more
 * passes, twice that for a perm that is not a permutation, plus 4 bytes per
 * column) must fit, else HX_ERR_INVALID_ARGUMENT.  Every call on a GPU shares
 * that GPU's array: calls that overlap in time on one GPU (other streams or
 * threads) race and must be ordered by the caller. */
#define HX_NTT_NOALLOC_SCRATCH_BYTES (64u << 20)
int hx_ntt_plan_exec_noalloc(const hx_ntt_plan* plan, const hx_view* a, const hx_view* result, hx_stream stream);
/* Measurement hook, no counterpart in the reference: the rate (butterflies per
 * second, whole GPU) of the NTT passes' 60-bit lazy butterfly with operands in
 * registers, i.e. the arithmetic ceiling of a pass of such butterflies on this
 * GPU.  Synchronises the stream. */
int hx_butterfly_rate(double* butterflies_per_s, hx_stream stream);
/* permute without the host synchronisation: perms must have been validated
 * (e.g. by an earlier hx_permute); out-of-range entries still never read
 * outside `a`, they write 0 and count in index_out_of_range. */
#define HX_PERMS_TRUSTED 1u
/* HX_PERMS_ROWS_SAME: every row of perms equals its row 0 (as hx_perms_rows_same
 * reports; a Galois map applied to every limb is such a table), so the gather
 * reads row 0 for every row: one row of indices stays in cache instead of the
 * whole [l, m] table being streamed.  The caller asserts it, like
 * HX_PERMS_TRUSTED; with a table whose rows differ the result is wrong. */
#define HX_PERMS_ROWS_SAME 2u
int hx_permute_ex(const hx_view* a, const hx_view* perms, const hx_view* result,
                  int64_t elementwise_axis, int64_t perm_axis, uint32_t flags, hx_stream stream);
/* *same = 1 when every row of the 2-D table perms equals its row 0, else 0.
 * Synchronises the stream. */
int hx_perms_rows_same(const hx_view* perms, int* same, hx_stream stream);

/* ------------------------------------------------- fused hybrid key-switch
 * No counterpart in the reference (whose API composes a key-switch from six
 * calls).  For B instances computes exactly
 *   out = permute(intt(axis_modsum(modmul_ttt(ntt(g_decomposition(x)), evk), axis 0)), perms)
 * with the views of that chain: x [B, m, k] int64, evk [dnum, B, 2, m, k],
 * perms [k, m] (each row a permutation), out [B, 2, m, k]; fwd and inv are plans
 * built from the same p and perm (inv with the inverse twiddles and m_inv).
 * 2^12 <= m <= 2^16.  The key is laid out once for the device by
 * hx_keyswitch_prepare_key into a caller-allocated buffer of
 * hx_keyswitch_key_bytes(plan, B) bytes; exec then runs three kernels and never
 * synchronises.  g_decomposition overflow is counted as by hx_g_decomposition.
 * The fused result equals the chain's (and the reference's) when every key word
 * is >= 0: prepare_key synchronises once and rejects a key with negative words
 * (HX_ERR_INVALID_ARGUMENT), whose chain result carries the reference's signed
 * INTT representatives; plan creation rejects tables outside [0, q), moduli
 * above 2^62 and digit sizes with 2^w - 1 + q > 2^63 - 1.  Such inputs belong to
 * the six-call chain.
 * Plan creation times a few launch shapes on scratch data (the split of log2 m
 * between the high and the low window, chosen once per (m, dnum) in a process,
 * and the limbs per inner workgroup, per (m, k, dnum)); every choice gives the
 * same result.  A key prepared by one plan serves every plan of the same m,
 * dnum, moduli prefix and NTT permutation in the process. */
typedef struct hx_keyswitch_plan hx_keyswitch_plan;
int hx_keyswitch_plan_create(hx_keyswitch_plan** out, const hx_ntt_plan* fwd, const hx_ntt_plan* inv,
                             const hx_view* perms, uint64_t dnum, uint64_t base_bits, hx_stream stream);
int hx_keyswitch_key_bytes(const hx_keyswitch_plan* plan, int64_t batch, size_t* bytes);
/* The tuned shape of a plan: low_bits = index bits transformed inside the
 * middle kernel (the rest in the first and last), limb_group = limbs per
 * middle-kernel workgroup, tile = sub-problems (columns) per workgroup of the
 * first and last kernels (32 or 16).  Any pointer may be NULL.  No counterpart
 * in the reference. */
int hx_keyswitch_plan_config(const hx_keyswitch_plan* plan, int32_t* low_bits, int32_t* limb_group, int32_t* tile);
int hx_keyswitch_prepare_key(const hx_keyswitch_plan* plan, const hx_view* evk, void* key, hx_stream stream);
int hx_keyswitch_exec(const hx_keyswitch_plan* plan, const hx_view* x, const void* key, size_t key_bytes,
                      const hx_view* out, hx_stream stream);
int hx_keyswitch_plan_destroy(hx_keyswitch_plan* plan);
/* hx_keyswitch_exec that also records events[0..3] (hipEvent_t, each may be
 * NULL) on `stream` before the first kernel and after each of the three, so a
 * caller can time every kernel of the launch sequence.  No counterpart in the
 * reference (measurement of the fused key-switch). */
int hx_keyswitch_exec_marked(const hx_keyswitch_plan* plan, const hx_view* x, const void* key, size_t key_bytes,
                             const hx_view* out, void* const* events, hx_stream stream);

/* hx_keyswitch_exec for serving many rotations with one plan per (m, k):
 *   key_limbs  limbs the key was prepared with (>= the plan's k; the first k
 *              are used), by a plan of the same m, moduli prefix and NTT perm;
 *   key_count  1: one key for every instance of the batch; else the batch size;
 *   row_map    int32 [m] device array from hx_keyswitch_row_map: the Galois
 *              permutation of every limb (NULL: the plan's own perms).
 * key_bytes must cover dnum * key_count * 2 * key_limbs * m * 8 bytes.
 * No counterpart in the reference (a vendor extension of the key-switch above). */
int hx_keyswitch_exec_ex(const hx_keyswitch_plan* plan, const hx_view* x, const void* key, size_t key_bytes,
                         int64_t key_limbs, int64_t key_count, const int32_t* row_map, const hx_view* out,
                         hx_stream stream);
/* hx_keyswitch_exec_ex of x (the second polynomial of each instance's
 * ciphertext ct) fused with one step of the diagonal matrix-vector product of
 * the serving loop: instead of writing the key-switch output KS, its last
 * kernel updates acc [B, 2, m, k] in place,
 *     acc_0[j] += pt[j] * (KS_0[j] + ct_0[perm[j]]),  acc_1[j] += pt[j] * KS_1[j]
 * (mod the plan's moduli, each step with the semantics of hx_rot_mac), where
 * perm is the Galois permutation row_map inverts and pt[j] = diags[b, diag mod
 * M, j mod M, t].  ct has acc's shape and strides; diags is [B, M, M, k],
 * contiguous per instance, M a power of two.  Equal to hx_keyswitch_exec_ex
 * followed by hx_rot_mac, without writing and re-reading KS.  No counterpart in
 * the reference (a vendor extension for serving). */
int hx_keyswitch_exec_rmac(const hx_keyswitch_plan* plan, const hx_view* x, const void* key, size_t key_bytes,
                           int64_t key_limbs, int64_t key_count, const int32_t* row_map, const hx_view* acc,
                           const hx_view* ct, const hx_view* diags, int64_t diag, hx_stream stream);
/* Exact negacyclic NTT of length 65536 for primes q < 2^28 on the INT8 matrix
 * cores (four-step: 256-point sub-transforms as 256 x 256 x 256 products of
 * 7-bit int8 planes with int32 accumulation, recombined and reduced mod q).
 * Equal to hx_ntt with twiddles[t, j] = psi_t^bitrev(j) and perm = bitrev.
 * The plan takes host tables (heal_amd/ntt_mfma.py builds them from q and psi):
 *   w2, w1  [k][4][16][4][64] x 16 bytes: the step matrices
 *           W2[u2, i2] = omega^(256 u2 i2) psi^(256 i2), W1[u1, i1] = omega^(256 u1 i1)
 *           (omega = psi^2), 7-bit planes in MFMA operand order;
 *   tmat    [k][256][256] uint32: T[u2, i1] = psi^i1 omega^(u2 i1);
 *   red     [k][10] uint64: q, floor((2^64 - 1) / q), 2^(7 s) mod q (s < 7), 0.
 * a, result: int64 [l, 65536, 1, k] views (limb-major layouts read columns
 * contiguously).  No counterpart in the reference. */
typedef struct hx_ntt_mfma_plan hx_ntt_mfma_plan;
int hx_ntt_mfma_plan_create(hx_ntt_mfma_plan** out, int64_t k, const void* w2, const void* w1, const void* tmat,
                            const void* red, hx_stream stream);
int hx_ntt_mfma_exec(const hx_ntt_mfma_plan* plan, const hx_view* a, const hx_view* result, hx_stream stream);
int hx_ntt_mfma_plan_destroy(hx_ntt_mfma_plan* plan);

/* Key-switches of mixed ring degree and limb count run together: the members
 * (plan i on x[i] with key[i], into out[i], exactly as hx_keyswitch_exec) are
 * fixed at creation and hx_keyswitch_group_exec runs all of them with one
 * launch per phase (K1, K2, K3) and window size, reading per-member launch
 * arguments from device memory.  Every member gets its own workspace, so
 * members may share a plan.  Results equal hx_keyswitch_exec of each member.
 * No counterpart in the reference. */
typedef struct hx_keyswitch_group hx_keyswitch_group;
int hx_keyswitch_group_create(hx_keyswitch_group** out, const hx_keyswitch_plan* const* plans, const hx_view* xs,
                              const void* const* keys, const size_t* key_bytes, const hx_view* outs, int count,
                              hx_stream stream);
int hx_keyswitch_group_exec(const hx_keyswitch_group* group, hx_stream stream);
int hx_keyswitch_group_destroy(hx_keyswitch_group* group);

/* row_map[perms[n]] = n for one permutation perms [m] (or [1, m]) of [0, m);
 * validates perms (synchronises the stream). */
int hx_keyswitch_row_map(const hx_keyswitch_plan* plan, const hx_view* perms, int32_t* row_map, hx_stream stream);

#ifdef __cplusplus
}
#endif
#endif /* HEAL_AMD_H */
