// hx_elementwise.hip — modmul/modsum/mod, axis_modsum, g_decomposition and
// permute for gfx950.
//
// All four are HBM-bound scans: one grid-stride loop per op over the
// coalesced iteration space of the output, walked in the memory order of the
// output (of the input for axis_modsum, which reads axis_len times more than
// it writes), 64-bit element offsets decoded with 32-bit fast division, and
// exact modular arithmetic from hx_arith.hpp.
// A modulus that varies only along the last axis (the RNS limb axis of HEAL
// tensors) is turned into a table of Möller–Granlund reciprocals in LDS once
// per workgroup, so the per-element cost is one double-width product and one
// reduction.
#include <algorithm>
#include <cstring>
#include <type_traits>

#include "hx_internal.hpp"

namespace hx {

void strided_copy(const hx_view* src, const hx_view* dst, hipStream_t st);

namespace {

constexpr int kBlock = 256;
constexpr int kTableMax = 512;

// How an operand supplies values to an elementwise kernel.
enum : int { SRC_SCALAR = 0, SRC_TABLE = 1, SRC_TENSOR = 2 };

struct ElemArgs {
    IterSpace it;          // operand order: 0 = result, 1 = a, 2 = b, 3 = p
    void* out;
    const void* a;
    const void* b;
    const void* p;
    int64_t a_scalar;
    int64_t b_scalar;
    int64_t p_scalar;
    int64_t p_table_stride;  // element stride of p along the last iteration dim (SRC_TABLE)
    int32_t p_table_len;
};

template <typename T>
__device__ __forceinline__ int64_t ld(const void* base, int64_t off) {
    return (int64_t)((const T*)base)[off];
}

template <typename T>
__device__ __forceinline__ void st_(void* base, int64_t off, int64_t v) {
    ((T*)base)[off] = (T)v;
}

template <typename T>
__device__ __forceinline__ int64_t to_t(int64_t v) {
    return (int64_t)(T)v;
}

// Modulus description of a signed value: |m| with reciprocal; M == 0 marks a
// zero modulus, whose result is defined here as 0.
__device__ __forceinline__ ModInfo modinfo_of(int64_t m) {
    uint64_t M = uabs64(m);
    if (M == 0) {
        ModInfo z;
        z.M = 0;
        z.d = 0;
        z.v = 0;
        z.s = 0;
        z.pad = 0;
        return z;
    }
    return make_modinfo(M);
}

// OP: 0 = modmul, 1 = modsum, 2 = remainder (a % b: the modulus is b, p unused),
//     3 = modular negation (the residue of -a in [0, |p|); b unused)
template <typename T, int OP, int BSRC, int PSRC>
__global__ __launch_bounds__(kBlock) void k_modop(ElemArgs A) {
    __shared__ ModInfo table[kTableMax];
    // modulus operand: p for OP 0/1, b for OP 2
    constexpr int MSRC = (OP == 2) ? BSRC : PSRC;
    ModInfo mconst;
    if (MSRC == SRC_SCALAR) mconst = modinfo_of(to_t<T>(OP == 2 ? A.b_scalar : A.p_scalar));
    if (MSRC == SRC_TABLE) {
        const void* mp = (OP == 2) ? A.b : A.p;
        for (int i = threadIdx.x; i < A.p_table_len; i += blockDim.x)
            table[i] = modinfo_of(ld<T>(mp, (int64_t)i * A.p_table_stride));
        __syncthreads();
    }
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < A.it.total; idx += stride) {
        int64_t off[4], last = 0;
        iter_offsets<4>(A.it, idx, off, &last);
        int64_t av = (OP == 2 && A.a == nullptr) ? to_t<T>(A.a_scalar) : ld<T>(A.a, off[1]);
        ModInfo mi;
        if (MSRC == SRC_SCALAR) mi = mconst;
        else if (MSRC == SRC_TABLE) mi = table[last];
        else mi = modinfo_of(ld<T>(OP == 2 ? A.b : A.p, OP == 2 ? off[2] : off[3]));
        int64_t r = 0;
        if (mi.M != 0) {
            if (OP == 2) {
                r = cpp_rem(mi, av);
            } else if (OP == 3) {
                const uint64_t ra = reduce64(mi, uabs64(av));
                r = (int64_t)(av < 0 ? ra : (ra ? mi.M - ra : 0));
            } else {
                int64_t bv = (BSRC == SRC_SCALAR) ? to_t<T>(A.b_scalar) : ld<T>(A.b, off[2]);
                r = (OP == 0) ? cpp_mulmod(mi, av, bv) : cpp_addmod(mi, av, bv);
            }
        }
        st_<T>(A.out, off[0], r);
    }
}

// Row fast path of modmul / modsum.  The last iteration dim (length k, the
// limb axis) has stride 1 in the result and in every tensor operand, and the
// modulus varies along it only (an LDS table).  A work-item takes a segment of
// kSeg consecutive elements of one row, so the row's offsets are decoded once
// per segment.  Operands that are residues (0 <= x < |p|) take a short path —
// one double-width product and one reduction, or one conditional subtraction
// — whose result is the residue the reference's `%` gives for them; any other
// element takes the reference's signed semantics as in k_modop.
// Segments are two int64 elements moved with one 16-byte access per operand
// (the host checks that every segment start is 16-byte aligned), so adjacent
// work-items touch adjacent 16-byte words.
constexpr int kSeg = 2;


template <typename T, int OP, int BSRC>
__global__ __launch_bounds__(kBlock) void k_modop_rows(ElemArgs A, FastDiv segs_fd, uint32_t segs, uint32_t nitems) {
    __shared__ ModInfo table[kTableMax];
    const int k = A.p_table_len;
    for (int i = threadIdx.x; i < k; i += blockDim.x) table[i] = modinfo_of(ld<T>(A.p, (int64_t)i * A.p_table_stride));
    __syncthreads();
    const IterSpace& it = A.it;
    for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < nitems; idx += gridDim.x * blockDim.x) {
        const uint32_t row = fdiv(idx, segs_fd), sg = idx - row * segs;
        int64_t o0 = 0, o1 = 0, o2 = 0;
        uint32_t n = row;
        for (int d = it.ndim - 2; d >= 0; --d) {
            const uint32_t q = fdiv(n, it.fd[d]);
            const int64_t c = (int64_t)(n - q * it.fd[d].d);
            o0 += c * it.strides[0][d];
            o1 += c * it.strides[1][d];
            o2 += c * it.strides[2][d];
            n = q;
        }
        const int t = (int)sg * kSeg;
        const longlong2 av = *(const longlong2*)((const int64_t*)A.a + o1 + t);
        longlong2 bv;
        if (BSRC == SRC_SCALAR) bv = make_longlong2(A.b_scalar, A.b_scalar);
        else bv = *(const longlong2*)((const int64_t*)A.b + o2 + t);
        longlong2 r;
        r.x = (int64_t)modop_fast(OP, table[t], av.x, bv.x);
        r.y = (int64_t)modop_fast(OP, table[t + 1], av.y, bv.y);
        __builtin_nontemporal_store(r.x, (int64_t*)A.out + o0 + t);
        __builtin_nontemporal_store(r.y, (int64_t*)A.out + o0 + t + 1);
    }
}

// Operand strides broadcast (right-aligned) onto the result's dims: size-1
// dims and missing leading dims get stride 0.  Operand dims larger than the
// result's (allowed for p by the reference's checks) index by the result's
// coordinate; smaller ones are out of range.
void broadcast_strides(const hx_view* v, const hx_view* r, int64_t* out, const char* name) {
    int shift = r->ndim - v->ndim;
    if (shift < 0) raise(HX_ERR_INVALID_ARGUMENT, std::string(name) + " has more dims than result.");
    for (int i = 0; i < r->ndim; ++i) out[i] = 0;
    for (int i = 0; i < v->ndim; ++i) {
        if (v->dims[i] == 1) continue;
        if (v->dims[i] < r->dims[i + shift] && view_numel(r) > 0) raise(HX_ERR_OUT_OF_RANGE, "Index out of bounds.");
        out[i + shift] = v->strides[i];
    }
}

// CHECK_DIMS_BROADCASTABLE (example_impl/modop_impl.cpp:89-101)
void check_broadcastable(const hx_view* t, const hx_view* r, const char* label) {
    if (t->ndim > r->ndim) raise(HX_ERR_INVALID_ARGUMENT, std::string(label) + " has more dims than result.");
    for (int i = 1; i <= t->ndim; ++i) {
        int64_t td = t->dims[t->ndim - i], rd = r->dims[r->ndim - i];
        if (td != 1 && td != rd)
            raise(HX_ERR_INVALID_ARGUMENT, std::string(label) + " not broadcast-compatible with result.");
    }
}

void require_int(const hx_view* v, const hx_view* r, const char* name) {
    if (v->dtype != HX_INT32 && v->dtype != HX_INT64)
        raise(HX_ERR_INVALID_ARGUMENT, std::string(name) + ": modular ops need int32 or int64 tensors");
    if (v->dtype != r->dtype) raise(HX_ERR_INVALID_ARGUMENT, std::string(name) + ": dtype differs from result");
}

template <typename T, int OP>
void launch_modop_t(ElemArgs& A, int bsrc, int psrc, hipStream_t st) {
    unsigned g = grid_for(A.it.total, kBlock);
#define HX_LAUNCH(B, P) hipLaunchKernelGGL((k_modop<T, OP, B, P>), dim3(g), dim3(kBlock), 0, st, A)
    if (bsrc == SRC_TENSOR) {
        if (psrc == SRC_SCALAR) HX_LAUNCH(SRC_TENSOR, SRC_SCALAR);
        else if (psrc == SRC_TABLE) HX_LAUNCH(SRC_TENSOR, SRC_TABLE);
        else HX_LAUNCH(SRC_TENSOR, SRC_TENSOR);
    } else if (bsrc == SRC_TABLE) {
        HX_LAUNCH(SRC_TABLE, SRC_SCALAR);
    } else {
        if (psrc == SRC_SCALAR) HX_LAUNCH(SRC_SCALAR, SRC_SCALAR);
        else if (psrc == SRC_TABLE) HX_LAUNCH(SRC_SCALAR, SRC_TABLE);
        else HX_LAUNCH(SRC_SCALAR, SRC_TENSOR);
    }
#undef HX_LAUNCH
    check_hip(hipGetLastError(), "modop launch");
}

// Decide whether the modulus operand (iteration operand index `mop`) varies
// only along the last iteration dim; if so use an LDS table.
int modulus_source(const ElemArgs& A, int mop) {
    const IterSpace& it = A.it;
    for (int d = 0; d < it.ndim - 1; ++d)
        if (it.strides[mop][d] != 0) return SRC_TENSOR;
    if (it.dims[it.ndim - 1] <= kTableMax) return SRC_TABLE;
    return SRC_TENSOR;
}

// modmul_* / modsum_* (example_impl/modop_impl.cpp:108-180)
void modop(int op, const hx_view* a, const hx_view* b, int64_t bs, const hx_view* p, int64_t ps,
           const hx_view* result, hipStream_t st) {
    check_view(a, "a");
    check_view(result, "result");
    if (b) check_view(b, "b");
    if (p) check_view(p, "p");
    require_int(result, result, "result");
    require_int(a, result, "a");
    if (b) require_int(b, result, "b");
    if (p) require_int(p, result, "p");
    check_broadcastable(a, result, "a");
    if (b) check_broadcastable(b, result, "b");
    if (p) {
        // CHECK_DIMS_MATCH_LAST (example_impl/modop_impl.cpp:84-87)
        if (p->ndim != 1 && (p->ndim == 0 || result->ndim == 0 || p->dims[p->ndim - 1] != result->dims[result->ndim - 1]))
            raise(HX_ERR_INVALID_ARGUMENT, "Last dimension of p must match last dimension of result.");
    }
    int64_t strides[kMaxOps][HX_MAX_DIMS] = {};
    for (int i = 0; i < result->ndim; ++i) strides[0][i] = result->strides[i];
    broadcast_strides(a, result, strides[1], "a");
    if (b) broadcast_strides(b, result, strides[2], "b");
    if (p) broadcast_strides(p, result, strides[3], "p");
    ElemArgs A;
    std::memset(&A, 0, sizeof A);
    A.it = make_iter(result->ndim, result->dims, 4, strides, p != nullptr);
    if (A.it.total == 0) return;
    A.out = result->data;
    A.a = a->data;
    A.b = b ? b->data : nullptr;
    A.p = p ? p->data : nullptr;
    A.b_scalar = bs;
    A.p_scalar = ps;
    int bsrc = b ? SRC_TENSOR : SRC_SCALAR;
    int psrc = SRC_SCALAR;
    if (p) {
        psrc = modulus_source(A, 3);
        if (psrc == SRC_TABLE) {
            A.p_table_stride = A.it.strides[3][A.it.ndim - 1];
            A.p_table_len = (int32_t)A.it.dims[A.it.ndim - 1];
        }
    }
    bool is32 = result->dtype == HX_INT32;
    // row fast path (see k_modop_rows)
    const int last = A.it.ndim - 1;
    const int64_t k = A.it.dims[last];
    const int64_t segs = (k + kSeg - 1) / kSeg;
    const int64_t nitems = A.it.total / (k ? k : 1) * segs;
    // every segment start must be 16-byte aligned: even row strides and k, aligned bases
    bool aligned = k % 2 == 0 && (uintptr_t)result->data % 16 == 0 && (uintptr_t)a->data % 16 == 0 &&
                   (!b || (uintptr_t)b->data % 16 == 0);
    for (int d = 0; d < last && aligned; ++d)
        aligned = A.it.strides[0][d] % 2 == 0 && A.it.strides[1][d] % 2 == 0 && (!b || A.it.strides[2][d] % 2 == 0);
    if (!is32 && aligned && (op == 0 || op == 1) && psrc == SRC_TABLE && A.it.use32 && A.it.strides[0][last] == 1 &&
        A.it.strides[1][last] == 1 && (!b || A.it.strides[2][last] == 1) && nitems < (int64_t(1) << 31)) {
        const FastDiv fd = make_fastdiv((uint32_t)segs);
        const unsigned g = grid_for(nitems, kBlock, 8192);
#define HX_ROWS(TT, OPV)                                                                                      \
    if (b) hipLaunchKernelGGL((k_modop_rows<TT, OPV, SRC_TENSOR>), dim3(g), dim3(kBlock), 0, st, A, fd,       \
                               (uint32_t)segs, (uint32_t)nitems);                                              \
    else hipLaunchKernelGGL((k_modop_rows<TT, OPV, SRC_SCALAR>), dim3(g), dim3(kBlock), 0, st, A, fd,         \
                            (uint32_t)segs, (uint32_t)nitems);
        if (op == 0) {
            HX_ROWS(int64_t, 0)
        } else {
            HX_ROWS(int64_t, 1)
        }
#undef HX_ROWS
        check_hip(hipGetLastError(), "modop rows launch");
        return;
    }
    if (op == 0) {
        if (is32) launch_modop_t<int32_t, 0>(A, bsrc, psrc, st);
        else launch_modop_t<int64_t, 0>(A, bsrc, psrc, st);
    } else if (op == 3) {
        if (is32) launch_modop_t<int32_t, 3>(A, SRC_SCALAR, psrc, st);
        else launch_modop_t<int64_t, 3>(A, SRC_SCALAR, psrc, st);
    } else {
        if (is32) launch_modop_t<int32_t, 1>(A, bsrc, psrc, st);
        else launch_modop_t<int64_t, 1>(A, bsrc, psrc, st);
    }
}

// mod_tt / mod_tc / mod_ct (example_impl/modop_impl.cpp:183-235)
void modrem(const hx_view* a, bool a_tensor, int64_t as, const hx_view* b, bool b_tensor, int64_t bs,
            const hx_view* result, hipStream_t st) {
    if (a_tensor && !a) raise(HX_ERR_INVALID_ARGUMENT, "a pointer must not be null.");
    if (b_tensor && !b) raise(HX_ERR_INVALID_ARGUMENT, "b pointer must not be null.");
    check_view(result, "result");
    require_int(result, result, "result");
    if (a) {
        check_view(a, "a");
        require_int(a, result, "a");
        if (!dims_equal(a, result)) raise(HX_ERR_INVALID_ARGUMENT, "a must have exactly the same shape as result.");
    }
    if (b) {
        check_view(b, "b");
        require_int(b, result, "b");
        if (!dims_equal(b, result)) raise(HX_ERR_INVALID_ARGUMENT, "b must have exactly the same shape as result.");
    }
    int64_t strides[kMaxOps][HX_MAX_DIMS] = {};
    for (int i = 0; i < result->ndim; ++i) {
        strides[0][i] = result->strides[i];
        if (a) strides[1][i] = a->strides[i];
        if (b) strides[2][i] = b->strides[i];
    }
    ElemArgs A;
    std::memset(&A, 0, sizeof A);
    A.it = make_iter(result->ndim, result->dims, 3, strides, false);
    if (A.it.total == 0) return;
    A.out = result->data;
    A.a = a ? a->data : nullptr;
    A.b = b ? b->data : nullptr;
    A.a_scalar = as;
    A.b_scalar = bs;
    int bsrc = b ? SRC_TENSOR : SRC_SCALAR;
    if (result->dtype == HX_INT32) launch_modop_t<int32_t, 2>(A, bsrc, SRC_SCALAR, st);
    else launch_modop_t<int64_t, 2>(A, bsrc, SRC_SCALAR, st);
}

// ------------------------------------------------------------ axis_modsum
struct AxisArgs {
    IterSpace it;  // over result: operand 0 = result, 1 = a (axis removed)
    void* out;
    const void* a;
    const void* p;
    int64_t axis_len;
    int64_t axis_stride;   // stride of a along the reduced axis
    int64_t p_stride;      // stride of p (1-D)
    int32_t table;         // 1: p indexed by the last coordinate from LDS
    int32_t k;
};

template <typename T>
__global__ __launch_bounds__(kBlock) void k_axis_modsum(AxisArgs A) {
    __shared__ ModInfo table[kTableMax];
    if (A.table) {
        for (int i = threadIdx.x; i < A.k; i += blockDim.x) table[i] = modinfo_of(ld<T>(A.p, (int64_t)i * A.p_stride));
        __syncthreads();
    }
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < A.it.total; idx += stride) {
        int64_t off[2], last = 0;
        iter_offsets<2>(A.it, idx, off, &last);
        ModInfo mi = A.table ? table[last] : modinfo_of(ld<T>(A.p, last * A.p_stride));
        int64_t sum = 0;
        const int64_t base = off[1];
        for (int64_t r = 0; r < A.axis_len; ++r) {
            int64_t s = (int64_t)((uint64_t)sum + (uint64_t)ld<T>(A.a, base + r * A.axis_stride));
            int64_t red;
            if (mi.M == 0) red = 0;
            else if (s >= 0 && (uint64_t)s < mi.M) red = s;
            else if (s >= 0 && mi.M <= (1ull << 62) && (uint64_t)s < 2 * mi.M) red = s - (int64_t)mi.M;
            else red = cpp_rem(mi, s);
            sum = to_t<T>(red);
        }
        st_<T>(A.out, off[0], sum);
    }
}

// axis_modsum (example_impl/axis_modsum_impl.cpp:12-76)
void axis_modsum(const hx_view* a, const hx_view* p, const hx_view* result, int64_t axis, hipStream_t st) {
    check_view(a, "a");
    check_view(p, "p");
    check_view(result, "result");
    require_int(result, result, "result");
    require_int(a, result, "a");
    require_int(p, result, "p");
    if (p->ndim != 1) raise(HX_ERR_INVALID_ARGUMENT, "p must be a 1D tensor of shape [k]");
    const int nd = a->ndim;
    if (axis < 0 || axis >= nd - 1)
        raise(HX_ERR_INVALID_ARGUMENT, "axis must be in range [0, ndim - 2] (can't reduce across last axis)");
    if (a->dims[nd - 1] != p->dims[0]) raise(HX_ERR_INVALID_ARGUMENT, "Last dimension of a must match shape of p");
    if (result->ndim != nd - 1) raise(HX_ERR_INVALID_ARGUMENT, "result must have the shape of a without axis");
    int64_t strides[kMaxOps][HX_MAX_DIMS] = {};
    for (int i = 0, j = 0; i < nd; ++i) {
        if (i == axis) continue;
        if (result->dims[j] != a->dims[i]) raise(HX_ERR_INVALID_ARGUMENT, "result must have the shape of a without axis");
        strides[0][j] = result->strides[j];
        strides[1][j] = a->strides[i];
        ++j;
    }
    AxisArgs A;
    std::memset(&A, 0, sizeof A);
    // walked in the memory order of a, which is axis_len times larger than the result
    A.it = make_iter(result->ndim, result->dims, 2, strides, true, 1);
    if (A.it.total == 0) return;
    A.out = result->data;
    A.a = a->data;
    A.p = p->data;
    A.axis_len = a->dims[axis];
    A.axis_stride = a->strides[axis];
    A.p_stride = p->strides[0];
    A.k = (int32_t)p->dims[0];
    A.table = p->dims[0] <= kTableMax;
    unsigned g = grid_for(A.it.total, kBlock);
    if (result->dtype == HX_INT32) hipLaunchKernelGGL(k_axis_modsum<int32_t>, dim3(g), dim3(kBlock), 0, st, A);
    else hipLaunchKernelGGL(k_axis_modsum<int64_t>, dim3(g), dim3(kBlock), 0, st, A);
    check_hip(hipGetLastError(), "axis_modsum launch");
}

// --------------------------------------------- modmul_axis_sum (fused)
// out = axis_modsum(modmul_ttt(a, b, p), p, axis): for every output element
// the products (T_DP(a) * T_DP(b)) % p along `axis` are summed with the
// reference's step semantics (sum = (sum + prod) % p, T arithmetic) without
// materialising the product tensor.
struct MasArgs {
    IterSpace it;  // over result: 0 = result, 1 = a, 2 = b (axis removed)
    void* out;
    const void* a;
    const void* b;
    const void* p;
    int64_t axis_len;
    int64_t a_axis_stride, b_axis_stride;
    int64_t p_stride;
    int32_t table;
    int32_t k;
};

template <typename T>
__global__ __launch_bounds__(kBlock) void k_modmul_axis_sum(MasArgs A) {
    __shared__ ModInfo table[kTableMax];
    if (A.table) {
        for (int i = threadIdx.x; i < A.k; i += blockDim.x) table[i] = modinfo_of(ld<T>(A.p, (int64_t)i * A.p_stride));
        __syncthreads();
    }
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < A.it.total; idx += stride) {
        int64_t off[3], last = 0;
        iter_offsets<3>(A.it, idx, off, &last);
        const ModInfo mi = A.table ? table[last] : modinfo_of(ld<T>(A.p, last * A.p_stride));
        int64_t sum = 0;
        if (mi.M != 0) {
            for (int64_t r = 0; r < A.axis_len; ++r) {
                const int64_t prod =
                    to_t<T>(cpp_mulmod(mi, ld<T>(A.a, off[1] + r * A.a_axis_stride), ld<T>(A.b, off[2] + r * A.b_axis_stride)));
                const int64_t s2 = (int64_t)((uint64_t)sum + (uint64_t)prod);
                int64_t red;
                if (s2 >= 0 && (uint64_t)s2 < mi.M) red = s2;
                else if (s2 >= 0 && mi.M <= (1ull << 62) && (uint64_t)s2 < 2 * mi.M) red = s2 - (int64_t)mi.M;
                else red = cpp_rem(mi, s2);
                sum = to_t<T>(red);
            }
        }
        st_<T>(A.out, off[0], sum);
    }
}

void modmul_axis_sum(const hx_view* a, const hx_view* b, const hx_view* p, const hx_view* result, int64_t axis,
                     hipStream_t st) {
    check_view(a, "a");
    check_view(b, "b");
    check_view(p, "p");
    check_view(result, "result");
    require_int(result, result, "result");
    require_int(a, result, "a");
    require_int(b, result, "b");
    require_int(p, result, "p");
    if (p->ndim != 1) raise(HX_ERR_INVALID_ARGUMENT, "p must be a 1D tensor of shape [k]");
    // full (broadcast) shape of a and b, right-aligned
    const int nd = std::max(a->ndim, b->ndim);
    int64_t full[HX_MAX_DIMS], sa[HX_MAX_DIMS] = {}, sb[HX_MAX_DIMS] = {};
    for (int i = 0; i < nd; ++i) {
        const int ia = i - (nd - a->ndim), ib = i - (nd - b->ndim);
        const int64_t da = ia >= 0 ? a->dims[ia] : 1, db = ib >= 0 ? b->dims[ib] : 1;
        if (da != db && da != 1 && db != 1) raise(HX_ERR_INVALID_ARGUMENT, "a and b are not broadcast-compatible");
        full[i] = std::max(da, db);
        sa[i] = (ia >= 0 && da != 1) ? a->strides[ia] : 0;
        sb[i] = (ib >= 0 && db != 1) ? b->strides[ib] : 0;
    }
    if (axis < 0 || axis >= nd - 1)
        raise(HX_ERR_INVALID_ARGUMENT, "axis must be in range [0, ndim - 2] (can't reduce across last axis)");
    if (full[nd - 1] != p->dims[0]) raise(HX_ERR_INVALID_ARGUMENT, "Last dimension of a must match shape of p");
    if (result->ndim != nd - 1) raise(HX_ERR_INVALID_ARGUMENT, "result must have the shape of a without axis");
    int64_t strides[kMaxOps][HX_MAX_DIMS] = {};
    for (int i = 0, j = 0; i < nd; ++i) {
        if (i == axis) continue;
        if (result->dims[j] != full[i]) raise(HX_ERR_INVALID_ARGUMENT, "result must have the shape of a without axis");
        strides[0][j] = result->strides[j];
        strides[1][j] = sa[i];
        strides[2][j] = sb[i];
        ++j;
    }
    MasArgs A;
    std::memset(&A, 0, sizeof A);
    A.it = make_iter(result->ndim, result->dims, 3, strides, true);
    if (A.it.total == 0) return;
    A.out = result->data;
    A.a = a->data;
    A.b = b->data;
    A.p = p->data;
    A.axis_len = full[axis];
    A.a_axis_stride = sa[axis];
    A.b_axis_stride = sb[axis];
    A.p_stride = p->strides[0];
    A.k = (int32_t)p->dims[0];
    A.table = p->dims[0] <= kTableMax;
    unsigned g = grid_for(A.it.total, kBlock);
    if (result->dtype == HX_INT32) hipLaunchKernelGGL(k_modmul_axis_sum<int32_t>, dim3(g), dim3(kBlock), 0, st, A);
    else hipLaunchKernelGGL(k_modmul_axis_sum<int64_t>, dim3(g), dim3(kBlock), 0, st, A);
    check_hip(hipGetLastError(), "modmul_axis_sum launch");
}

// --------------------------------------------- diagonal rotate-multiply-accumulate
// One step of the diagonal matrix-vector product on ciphertexts, the calls
//     rot_0 = modsum_ttt(ks_0, permute(ct_0, perm))   (rot = ct when ks is null)
//     rot_1 = ks_1
//     acc_c = modsum_ttt(acc_c, modmul_ttt(rot_c, pt))
// with pt[b, j, t] = diags[b, diag, j mod M, t] (the matrix stored by
// diagonals: diags[b, r, i, t] = M_b[i, (i + r) mod M] mod q_t, so that one
// step reads one contiguous [M, L] block per instance), fused: each
// work-item takes one (b, j, t) for both polynomials c.  Every step is the
// reference op's own (modop_fast: the residue fast path, the C++ semantics
// otherwise), so the result equals the five calls.
struct RotMacArgs {
    int64_t* acc;
    const int64_t* ks;
    const int64_t* ct;
    const int64_t* perm;
    const int64_t* diags;
    const int64_t* p;
    int64_t acc_b, ks_b, ct_b, diags_b;  // batch strides
    int64_t n, L, M, diag;
    int64_t total;                      // B * n * L
    FastDiv fd_L, fd_n;
    DevStatus* status;
};

__global__ __launch_bounds__(kBlock) void k_rot_mac(RotMacArgs A) {
    __shared__ ModInfo table[kTableMax];
    for (int i = threadIdx.x; i < A.L; i += blockDim.x) table[i] = modinfo_of(A.p[i]);
    __syncthreads();
    const int64_t plane = A.n * A.L;
    unsigned long long bad = 0;
    for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < (uint32_t)A.total;
         idx += gridDim.x * blockDim.x) {
        const uint32_t rest = fdiv(idx, A.fd_L), t = idx - rest * (uint32_t)A.L;
        const uint32_t b = fdiv(rest, A.fd_n), j = rest - b * (uint32_t)A.n;
        const ModInfo mi = table[t];
        const int64_t e = (int64_t)j * A.L + t;
        const int64_t pt = A.diags[b * A.diags_b + (A.diag * A.M + (j & (A.M - 1))) * A.L + t];
        int64_t r0, r1;
        if (A.ks) {
            const int64_t src = A.perm[j];
            const bool ok = src >= 0 && src < A.n;
            bad += !ok;
            const int64_t c0 = ok ? A.ct[b * A.ct_b + src * A.L + t] : 0;
            r0 = (int64_t)modop_fast(1, mi, A.ks[b * A.ks_b + e], c0);
            r1 = A.ks[b * A.ks_b + plane + e];
        } else {
            r0 = A.ct[b * A.ct_b + e];
            r1 = A.ct[b * A.ct_b + plane + e];
        }
        int64_t* acc = A.acc + b * A.acc_b + e;
        acc[0] = (int64_t)modop_fast(1, mi, acc[0], (int64_t)modop_fast(0, mi, r0, pt));
        acc[plane] = (int64_t)modop_fast(1, mi, acc[plane], (int64_t)modop_fast(0, mi, r1, pt));
    }
    if (bad) atomicAdd(&A.status->index_out_of_range, bad);
}

// [B, 2, n, L] (or [B, M, M, L] for diags) with unit strides inside an instance
void require_instances(const hx_view* v, const int64_t* dims, const char* name) {
    check_view(v, name);
    if (v->dtype != HX_INT64) raise(HX_ERR_INVALID_ARGUMENT, std::string(name) + " must be int64");
    if (v->ndim != 4) raise(HX_ERR_INVALID_ARGUMENT, std::string(name) + " must be 4-D");
    int64_t s = 1;
    for (int d = 3; d >= 0; --d) {
        if (v->dims[d] != dims[d]) raise(HX_ERR_INVALID_ARGUMENT, std::string(name) + " has the wrong shape");
        if (d > 0 && v->dims[d] > 1 && v->strides[d] != s)
            raise(HX_ERR_INVALID_ARGUMENT, std::string(name) + " must be contiguous within an instance");
        s *= v->dims[d];
    }
}

void rot_mac(const hx_view* acc, const hx_view* ks, const hx_view* ct, const hx_view* perm, const hx_view* diags,
             int64_t diag, const hx_view* p, hipStream_t st) {
    check_view(acc, "acc");
    check_view(p, "p");
    if (acc->ndim != 4) raise(HX_ERR_INVALID_ARGUMENT, "acc must have shape [B, 2, n, L]");
    const int64_t B = acc->dims[0], n = acc->dims[2], L = acc->dims[3];
    const int64_t shape[4] = {B, 2, n, L};
    require_instances(acc, shape, "acc");
    require_instances(ct, shape, "ct");
    if (ks) require_instances(ks, shape, "ks");
    if (!diags || diags->ndim != 4) raise(HX_ERR_INVALID_ARGUMENT, "diags must have shape [B, M, M, L]");
    const int64_t M = diags->dims[1];
    const int64_t mshape[4] = {B, M, M, L};
    require_instances(diags, mshape, "diags");
    if (M <= 0 || (M & (M - 1)) != 0) raise(HX_ERR_INVALID_ARGUMENT, "the matrix size M must be a power of two");
    if (p->dtype != HX_INT64 || p->ndim != 1 || p->dims[0] != L || (L > 1 && p->strides[0] != 1))
        raise(HX_ERR_INVALID_ARGUMENT, "p must be a contiguous int64 tensor of shape [L]");
    if (L > kTableMax) raise(HX_ERR_INVALID_ARGUMENT, "at most " + std::to_string(kTableMax) + " limbs");
    if (ks) {
        if (!perm || perm->dtype != HX_INT64 || perm->ndim != 1 || perm->dims[0] != n ||
            (n > 1 && perm->strides[0] != 1))
            raise(HX_ERR_INVALID_ARGUMENT, "perm must be a contiguous int64 tensor of shape [n]");
    }
    if (diag < 0) raise(HX_ERR_INVALID_ARGUMENT, "diag must be >= 0");
    const int64_t total = B * n * L;
    if (total == 0) return;
    if (total >= (int64_t(1) << 31)) raise(HX_ERR_INVALID_ARGUMENT, "rot_mac: B * n * L must be below 2^31");
    if (views_overlap(acc, ct) || (ks && views_overlap(acc, ks)) || views_overlap(acc, diags))
        raise(HX_ERR_INVALID_ARGUMENT, "acc must not overlap the other operands");
    RotMacArgs A;
    std::memset(&A, 0, sizeof A);
    A.acc = (int64_t*)acc->data;
    A.ks = ks ? (const int64_t*)ks->data : nullptr;
    A.ct = (const int64_t*)ct->data;
    A.perm = ks ? (const int64_t*)perm->data : nullptr;
    A.diags = (const int64_t*)diags->data;
    A.p = (const int64_t*)p->data;
    A.acc_b = acc->strides[0];
    A.ks_b = ks ? ks->strides[0] : 0;
    A.ct_b = ct->strides[0];
    A.diags_b = diags->strides[0];
    A.n = n;
    A.L = L;
    A.M = M;
    A.diag = diag % M;
    A.total = total;
    A.fd_L = make_fastdiv((uint32_t)L);
    A.fd_n = make_fastdiv((uint32_t)n);
    A.status = device_status();
    hipLaunchKernelGGL(k_rot_mac, dim3(grid_for(total, kBlock, 8192)), dim3(kBlock), 0, st, A);
    check_hip(hipGetLastError(), "rot_mac launch");
}

// ------------------------------------------------------------ abs, fill
template <typename T>
__global__ __launch_bounds__(kBlock) void k_abs(ElemArgs A) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < A.it.total; idx += stride) {
        int64_t off[2], last = 0;
        iter_offsets<2>(A.it, idx, off, &last);
        const T v = ((const T*)A.a)[off[1]];
        ((T*)A.out)[off[0]] = v < 0 ? (T)(0 - (typename std::make_unsigned<T>::type)v) : v;  // wraps at the minimum
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(ElemArgs A, T value) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < A.it.total; idx += stride) {
        int64_t off[1], last = 0;
        iter_offsets<1>(A.it, idx, off, &last);
        ((T*)A.out)[off[0]] = value;
    }
}

void abs_values(const hx_view* a, const hx_view* result, hipStream_t st) {
    check_view(a, "a");
    check_view(result, "result");
    require_int(result, result, "result");
    require_int(a, result, "a");
    if (!dims_equal(a, result)) raise(HX_ERR_INVALID_ARGUMENT, "a must have exactly the same shape as result.");
    int64_t strides[kMaxOps][HX_MAX_DIMS] = {};
    for (int i = 0; i < result->ndim; ++i) {
        strides[0][i] = result->strides[i];
        strides[1][i] = a->strides[i];
    }
    ElemArgs A;
    std::memset(&A, 0, sizeof A);
    A.it = make_iter(result->ndim, result->dims, 2, strides, false);
    if (A.it.total == 0) return;
    A.out = result->data;
    A.a = a->data;
    unsigned g = grid_for(A.it.total, kBlock);
    if (result->dtype == HX_INT32) hipLaunchKernelGGL(k_abs<int32_t>, dim3(g), dim3(kBlock), 0, st, A);
    else hipLaunchKernelGGL(k_abs<int64_t>, dim3(g), dim3(kBlock), 0, st, A);
    check_hip(hipGetLastError(), "abs launch");
}

void set_const(const hx_view* a, int64_t ivalue, double fvalue, hipStream_t st) {
    check_view(a, "a");
    int64_t strides[kMaxOps][HX_MAX_DIMS] = {};
    for (int i = 0; i < a->ndim; ++i) strides[0][i] = a->strides[i];
    ElemArgs A;
    std::memset(&A, 0, sizeof A);
    A.it = make_iter(a->ndim, a->dims, 1, strides, false);
    if (A.it.total == 0) return;
    A.out = a->data;
    unsigned g = grid_for(A.it.total, kBlock);
    if (a->dtype == HX_INT32) hipLaunchKernelGGL(k_fill<int32_t>, dim3(g), dim3(kBlock), 0, st, A, (int32_t)ivalue);
    else if (a->dtype == HX_INT64) hipLaunchKernelGGL(k_fill<int64_t>, dim3(g), dim3(kBlock), 0, st, A, ivalue);
    else hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(kBlock), 0, st, A, fvalue);
    check_hip(hipGetLastError(), "fill launch");
}

// -------------------------------------------------------- g_decomposition
struct GdArgs {
    IterSpace it;  // over a: operand 0 = result (digit 0), 1 = a, 2 = row-major index of the element
    void* out;
    const void* a;
    int64_t digit_stride;
    uint32_t power;
    uint32_t base_bits;
    DevStatus* status;
    unsigned long long tag;  // the call's tag in the overflow log
};

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gdecomp(GdArgs A) {
    const uint64_t mask = (A.base_bits == 0) ? 0 : ((~0ull) >> (64 - A.base_bits));
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    unsigned long long over = 0;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < A.it.total; idx += stride) {
        int64_t off[3], last;
        iter_offsets<3>(A.it, idx, off, &last);
        int64_t v = ld<T>(A.a, off[1]);
        for (uint32_t d = 0; d < A.power; ++d) {
            uint64_t u = (uint64_t)v;  // size_t arithmetic of the reference
            st_<T>(A.out, off[0] + (int64_t)d * A.digit_stride, to_t<T>((int64_t)(u & mask)));
            v = to_t<T>((int64_t)(A.base_bits ? (u >> A.base_bits) : u));
        }
        if (v > 0) {
            if (!over) gd_log_element(A.status, A.tag, (unsigned long long)off[2]);
            ++over;
        }
    }
    if (over) atomicAdd(&A.status->gdecomp_overflow, over);
}

// One digit per work-item, the digit index fastest: when the digit axis is
// contiguous, adjacent work-items write adjacent words, so one store
// instruction covers whole lines (one work-item per element writes `power`
// words a digit-stride apart, which splits every store over many lines).  The
// work-items of one element each load it (a cache hit after the first) and
// run the reference's digit loop up to their digit: every shift truncates to T.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gdecomp_digits(GdArgs A, FastDiv pw) {
    const uint64_t mask = (A.base_bits == 0) ? 0 : ((~0ull) >> (64 - A.base_bits));
    const uint32_t n = (uint32_t)(A.it.total * A.power);
    unsigned long long over = 0;
    for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += gridDim.x * blockDim.x) {
        const uint32_t e = fdiv(idx, pw), d = idx - e * A.power;
        int64_t off[3], last;
        iter_offsets<3>(A.it, (int64_t)e, off, &last);
        int64_t v = ld<T>(A.a, off[1]);
        for (uint32_t i = 0; i < d; ++i) {
            const uint64_t u = (uint64_t)v;  // size_t arithmetic of the reference
            v = to_t<T>((int64_t)(A.base_bits ? (u >> A.base_bits) : u));
        }
        const uint64_t u = (uint64_t)v;
        st_<T>(A.out, off[0] + d, to_t<T>((int64_t)(u & mask)));
        if (d + 1 == A.power && to_t<T>((int64_t)(A.base_bits ? (u >> A.base_bits) : u)) > 0) {
            if (!over) gd_log_element(A.status, A.tag, (unsigned long long)off[2]);
            ++over;
        }
    }
    if (over) atomicAdd(&A.status->gdecomp_overflow, over);
}

// Contiguous a and result (digit axis innermost): a workgroup takes 256
// consecutive elements, one per work-item, runs the reference's digit loop
// into LDS and stores the 256 * power digits as one contiguous run, so that
// both the loads and the stores are whole lines and every element is read once.
constexpr int kGdFlatMaxPower = 8;

template <typename T>
__global__ __launch_bounds__(256) void k_gdecomp_flat(GdArgs A, int64_t n) {
    __shared__ T buf[256 * kGdFlatMaxPower];
    const uint64_t mask = (A.base_bits == 0) ? 0 : ((~0ull) >> (64 - A.base_bits));
    const int P = (int)A.power;
    const T* a = (const T*)A.a;
    T* out = (T*)A.out;
    unsigned long long over = 0;
    for (int64_t base = (int64_t)blockIdx.x * 256; base < n; base += (int64_t)gridDim.x * 256) {
        const int cnt = (int)min<int64_t>(256, n - base);
        if ((int)threadIdx.x < cnt) {
            int64_t v = a[base + threadIdx.x];
            for (int d = 0; d < P; ++d) {
                const uint64_t u = (uint64_t)v;  // size_t arithmetic of the reference
                buf[threadIdx.x * P + d] = (T)to_t<T>((int64_t)(u & mask));
                v = to_t<T>((int64_t)(A.base_bits ? (u >> A.base_bits) : u));
            }
            if (v > 0) {
                if (!over) gd_log_element(A.status, A.tag, (unsigned long long)(base + threadIdx.x));
                ++over;
            }
        }
        __syncthreads();
        T* dst = out + base * P;
        for (int i = threadIdx.x; i < cnt * P; i += 256) __builtin_nontemporal_store(buf[i], dst + i);
        __syncthreads();
    }
    if (over) atomicAdd(&A.status->gdecomp_overflow, over);
}

// g_decomposition (example_impl/g_decomposition_impl.cpp:10-67)
void g_decomposition(const hx_view* a, const hx_view* result, uint64_t power, uint64_t base_bits,
                     unsigned long long tag, hipStream_t st) {
    check_view(a, "a");
    check_view(result, "result");
    require_int(result, result, "result");
    require_int(a, result, "a");
    bool ok = result->ndim == a->ndim + 1 && result->dims[result->ndim - 1] == (int64_t)power;
    for (int i = 0; ok && i < a->ndim; ++i) ok = result->dims[i] == a->dims[i];
    if (!ok) raise(HX_ERR_INVALID_ARGUMENT, "Output must have shape a.shape + [power]");
    if (base_bits >= 64) raise(HX_ERR_INVALID_ARGUMENT, "base_bits must be < 64");
    if (power > 0xffffffffull) raise(HX_ERR_INVALID_ARGUMENT, "power too large");
    int64_t strides[kMaxOps][HX_MAX_DIMS] = {};
    int64_t rm = 1;  // row-major strides of a's shape: the element's index for the overflow log
    for (int i = a->ndim - 1; i >= 0; --i) {
        strides[0][i] = result->strides[i];
        strides[1][i] = a->strides[i];
        strides[2][i] = rm;
        rm *= a->dims[i];
    }
    GdArgs A;
    std::memset(&A, 0, sizeof A);
    A.it = make_iter(a->ndim, a->dims, 3, strides, false);
    if (A.it.total == 0) return;
    A.out = result->data;
    A.a = a->data;
    A.digit_stride = result->strides[result->ndim - 1];
    A.power = (uint32_t)power;
    A.base_bits = (uint32_t)base_bits;
    A.status = device_status();
    A.tag = tag;
    const int64_t nout = A.it.total * (int64_t)power;
    const bool flat = A.it.ndim == 1 && A.it.strides[1][0] == 1 && A.it.strides[2][0] == 1 &&
                      A.it.strides[0][0] == (int64_t)power;
    if (flat && A.digit_stride == 1 && power >= 1 && power <= (uint64_t)kGdFlatMaxPower) {
        const unsigned g = grid_for(A.it.total, 256, 16384);
        if (result->dtype == HX_INT32)
            hipLaunchKernelGGL(k_gdecomp_flat<int32_t>, dim3(g), dim3(256), 0, st, A, A.it.total);
        else hipLaunchKernelGGL(k_gdecomp_flat<int64_t>, dim3(g), dim3(256), 0, st, A, A.it.total);
    } else if (A.digit_stride == 1 && power >= 1 && nout < (int64_t(1) << 31)) {
        const FastDiv pw = make_fastdiv((uint32_t)power);
        unsigned g = grid_for(nout, kBlock);
        if (result->dtype == HX_INT32) hipLaunchKernelGGL(k_gdecomp_digits<int32_t>, dim3(g), dim3(kBlock), 0, st, A, pw);
        else hipLaunchKernelGGL(k_gdecomp_digits<int64_t>, dim3(g), dim3(kBlock), 0, st, A, pw);
    } else {
        unsigned g = grid_for(A.it.total, kBlock);
        if (result->dtype == HX_INT32) hipLaunchKernelGGL(k_gdecomp<int32_t>, dim3(g), dim3(kBlock), 0, st, A);
        else hipLaunchKernelGGL(k_gdecomp<int64_t>, dim3(g), dim3(kBlock), 0, st, A);
    }
    check_hip(hipGetLastError(), "g_decomposition launch");
}

// ------------------------------------------------------------------ permute
struct PermArgs {
    IterSpace it;  // over result dims (no merging across the two axes): 0 = result, 1 = a
    void* out;
    const void* a;
    const void* perms;
    int32_t ew_dim;   // iteration dim index of elementwise_axis
    int32_t pa_dim;   // iteration dim index of perm_axis
    int64_t a_pa_stride;
    int64_t ps0, ps1;  // perms strides
    int64_t m;
    DevStatus* status;
};

template <typename T>
__global__ __launch_bounds__(kBlock) void k_permute(PermArgs A) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    unsigned long long bad = 0;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < A.it.total; idx += stride) {
        // full decode (the two permutation axes are kept as separate dims)
        int64_t ooff = 0, aoff = 0, cew = 0, cpa = 0;
        int64_t n = idx;
        for (int d = A.it.ndim - 1; d >= 0; --d) {
            int64_t c;
            if (A.it.use32) {
                uint32_t q = fdiv((uint32_t)n, A.it.fd[d]);
                c = (uint32_t)n - q * A.it.fd[d].d;
                n = q;
            } else {
                c = n % A.it.dims[d];
                n /= A.it.dims[d];
            }
            if (d == A.ew_dim) cew = c;
            if (d == A.pa_dim) cpa = c;
            ooff += c * A.it.strides[0][d];
            aoff += c * A.it.strides[1][d];
        }
        int64_t pi = ld<T>(A.perms, cew * A.ps0 + cpa * A.ps1);
        int64_t v = 0;
        if (pi >= 0 && pi < A.m) v = ld<T>(A.a, aoff + (pi - cpa) * A.a_pa_stride);
        else ++bad;
        st_<T>(A.out, ooff, v);
    }
    if (bad) atomicAdd(&A.status->index_out_of_range, bad);
}

// Fast path of permute for a validated table whose rows all equal row 0 (one
// map for every row, as for a Galois automorphism of every limb) when the
// elementwise axis is the last dim with stride 1 in both a and the result: a
// work-item takes kSeg consecutive elements along it, decodes its offsets and
// loads the permutation entry once, and copies the segment from the source
// row with one 16-byte access (int64, segment starts 16-byte aligned).
// Entries outside [0, m) write 0 and are counted, as in k_permute.
__global__ __launch_bounds__(kBlock) void k_permute_rows(PermArgs A, FastDiv segs_fd, uint32_t segs, uint32_t nitems) {
    const IterSpace& it = A.it;
    unsigned long long bad = 0;
    for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < nitems; idx += gridDim.x * blockDim.x) {
        const uint32_t row = fdiv(idx, segs_fd), sg = idx - row * segs;
        int64_t ooff = 0, aoff = 0, cpa = 0;
        uint32_t n = row;
        for (int d = it.ndim - 2; d >= 0; --d) {
            const uint32_t q = fdiv(n, it.fd[d]);
            const int64_t c = (int64_t)(n - q * it.fd[d].d);
            if (d == A.pa_dim) cpa = c;
            ooff += c * it.strides[0][d];
            aoff += c * it.strides[1][d];
            n = q;
        }
        const int64_t pi = ld<int64_t>(A.perms, cpa * A.ps1);
        const bool ok = pi >= 0 && pi < A.m;
        const int64_t src = aoff + (pi - cpa) * A.a_pa_stride;
        const int t = (int)sg * kSeg;
        longlong2 v = make_longlong2(0, 0);
        if (ok) v = *(const longlong2*)((const int64_t*)A.a + src + t);
        else bad += 2;
        __builtin_nontemporal_store(v.x, (int64_t*)A.out + ooff + t);
        __builtin_nontemporal_store(v.y, (int64_t*)A.out + ooff + t + 1);
    }
    if (bad) atomicAdd(&A.status->index_out_of_range, bad);
}

// Entries of rows 1.. of a 2-D table that differ from the same column of row 0.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_perm_rows_diff(const void* perms, int64_t l, int64_t m, int64_t s0,
                                                           int64_t s1, unsigned long long* count) {
    const int64_t total = (l - 1) * m;
    unsigned long long diff = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = 1 + i / m, c = i % m;
        diff += ld<T>(perms, r * s0 + c * s1) != ld<T>(perms, c * s1);
    }
    if (diff) atomicAdd(count, diff);
}

// entries of an [l, m] table outside [0, m)
template <typename T>
__global__ void k_perm_check(const void* perms, int64_t l, int64_t m, int64_t s0, int64_t s1,
                             unsigned long long* count) {
    const int64_t total = l * m;
    unsigned long long bad = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = ld<T>(perms, (i / m) * s0 + (i % m) * s1);
        bad += v < 0 || v >= m;
    }
    if (bad) atomicAdd(count, bad);
}

}  // namespace

// Counts the entries of a 2-D table with one kernel (rows_diff: entries that
// differ from row 0; else entries outside [0, m)) and waits for the count.
// Runs on the side stream while st is being captured.
static unsigned long long perms_count(const hx_view* perms, bool rows_diff, hipStream_t caller) {
    const int64_t l = perms->dims[0], m = perms->dims[1];
    SyncScope sc(caller);
    hipStream_t st = sc.stream();
    unsigned long long* count = nullptr;
    check_hip(hipMallocAsync((void**)&count, 256, st), "hipMallocAsync(table count)");
    unsigned long long h = 0;
    hipError_t e = hipSuccess;
    fill_zero(count, 256, st);
    const int64_t n = rows_diff ? (l - 1) * m : l * m;
    const unsigned g = grid_for(n, kBlock);
    const bool i32 = perms->dtype == HX_INT32;
    if (rows_diff) {
        if (i32)
            hipLaunchKernelGGL(k_perm_rows_diff<int32_t>, dim3(g), dim3(kBlock), 0, st, perms->data, l, m,
                               perms->strides[0], perms->strides[1], count);
        else
            hipLaunchKernelGGL(k_perm_rows_diff<int64_t>, dim3(g), dim3(kBlock), 0, st, perms->data, l, m,
                               perms->strides[0], perms->strides[1], count);
    } else {
        if (i32)
            hipLaunchKernelGGL(k_perm_check<int32_t>, dim3(g), dim3(kBlock), 0, st, perms->data, l, m,
                               perms->strides[0], perms->strides[1], count);
        else
            hipLaunchKernelGGL(k_perm_check<int64_t>, dim3(g), dim3(kBlock), 0, st, perms->data, l, m,
                               perms->strides[0], perms->strides[1], count);
    }
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&h, count, sizeof h, hipMemcpyDeviceToHost, st);
    (void)hipFreeAsync(count, st);
    check_hip(e, "table check");
    sc.sync();
    return h;
}

bool perms_rows_same(const hx_view* perms, hipStream_t st) {
    check_view(perms, "perms");
    if (perms->dtype != HX_INT32 && perms->dtype != HX_INT64)
        raise(HX_ERR_INVALID_ARGUMENT, "perms: an int32 or int64 table is expected");
    if (perms->ndim != 2) raise(HX_ERR_INVALID_ARGUMENT, "Perms must be a 2-D table [l, m].");
    const int64_t l = perms->dims[0], m = perms->dims[1];
    if (l <= 1 || m == 0) return true;
    return perms_count(perms, true, st) == 0;
}

// permute (example_impl/permute_impl.cpp:8-65).  rows_same: every row of perms
// equals row 0 (asserted by the caller), so the gather reads row 0 only.
void permute(const hx_view* a, const hx_view* perms, const hx_view* result, int64_t ew, int64_t pa, bool trusted,
             bool rows_same, hipStream_t st) {
    check_view(a, "a");
    check_view(perms, "perms");
    check_view(result, "result");
    require_int(result, result, "result");
    require_int(a, result, "a");
    require_int(perms, result, "perms");
    const int nd = a->ndim;
    if (!dims_equal(result, a)) raise(HX_ERR_INVALID_ARGUMENT, "Result shape must match input shape.");
    if (ew < 0 || ew >= nd || pa < 0 || pa >= nd) raise(HX_ERR_INVALID_ARGUMENT, "Axis out of bounds.");
    if (ew == pa) raise(HX_ERR_INVALID_ARGUMENT, "elementwise_axis and perm_axis must be different.");
    const int64_t l = a->dims[ew], m = a->dims[pa];
    if (perms->ndim != 2 || perms->dims[0] != l || perms->dims[1] != m)
        raise(HX_ERR_INVALID_ARGUMENT,
              "Perms must have shape [l, m] where l and m match a.shape at elementwise and perm axes.");
    if (view_numel(a) == 0) return;
    // the table is checked before anything is written (synchronous)
    if (!trusted && perms_count(perms, false, st)) raise(HX_ERR_OUT_OF_RANGE, "Permutation index out of bounds.");
    // A result that overlaps the input is computed from a private copy of the input.
    hx_view src = *a;
    void* tmp = nullptr;
    if (views_overlap(a, result)) {
        size_t bytes = (size_t)view_numel(a) * dtype_size(a->dtype);
        check_hip(hipMallocAsync(&tmp, bytes, st), "hipMallocAsync(permute tmp)");
        src.data = tmp;
        int64_t s = 1;
        for (int i = nd - 1; i >= 0; --i) {
            src.strides[i] = s;
            s *= a->dims[i];
        }
        strided_copy(a, &src, st);
    }
    DevStatus* status = device_status();
    // Iteration space without size-1 dropping or merging, so that the two
    // permutation axes keep their own coordinates.
    PermArgs A;
    std::memset(&A, 0, sizeof A);
    IterSpace& it = A.it;
    it.ndim = nd;
    it.total = view_numel(a);
    it.use32 = it.total < (int64_t(1) << 31);
    for (int i = 0; i < nd; ++i) {
        it.dims[i] = a->dims[i];
        it.fd[i] = make_fastdiv(it.use32 ? (uint32_t)a->dims[i] : 1u);
        it.strides[0][i] = result->strides[i];
        it.strides[1][i] = src.strides[i];
    }
    A.out = result->data;
    A.a = src.data;
    A.perms = perms->data;
    A.ew_dim = (int32_t)ew;
    A.pa_dim = (int32_t)pa;
    A.a_pa_stride = src.strides[pa];
    A.ps0 = rows_same ? 0 : perms->strides[0];
    A.ps1 = perms->strides[1];
    A.m = m;
    A.status = status;
    // row fast path (see k_permute_rows)
    const int last = nd - 1;
    const int64_t k = a->dims[last];
    const int64_t segs = (k + kSeg - 1) / kSeg;
    const int64_t nitems = k ? it.total / k * segs : 0;
    bool aligned = result->dtype == HX_INT64 && k % 2 == 0 && (uintptr_t)result->data % 16 == 0 &&
                   (uintptr_t)src.data % 16 == 0;
    for (int d = 0; d < last && aligned; ++d) aligned = result->strides[d] % 2 == 0 && src.strides[d] % 2 == 0;
    if (aligned && rows_same && ew == last && pa != last && it.use32 && result->strides[last] == 1 &&
        src.strides[last] == 1 && nitems < (int64_t(1) << 31)) {
        const FastDiv fd = make_fastdiv((uint32_t)segs);
        const unsigned g = grid_for(nitems, kBlock, 8192);
        hipLaunchKernelGGL(k_permute_rows, dim3(g), dim3(kBlock), 0, st, A, fd, (uint32_t)segs, (uint32_t)nitems);
    } else {
        unsigned g = grid_for(it.total, kBlock);
        if (result->dtype == HX_INT32) hipLaunchKernelGGL(k_permute<int32_t>, dim3(g), dim3(kBlock), 0, st, A);
        else hipLaunchKernelGGL(k_permute<int64_t>, dim3(g), dim3(kBlock), 0, st, A);
    }
    check_hip(hipGetLastError(), "permute launch");
    if (tmp) check_hip(hipFreeAsync(tmp, st), "hipFreeAsync");
}

}  // namespace hx

using namespace hx;

extern "C" {

int hx_modmul_ttt(const hx_view* a, const hx_view* b, const hx_view* p, const hx_view* result, hx_stream stream) {
    HX_API_BEGIN
    if (!b) raise(HX_ERR_INVALID_ARGUMENT, "b pointer must not be null.");
    if (!p) raise(HX_ERR_INVALID_ARGUMENT, "p pointer must not be null.");
    hx::modop(0, a, b, 0, p, 0, result, as_stream(stream));
    HX_API_END
}
int hx_modmul_ttc(const hx_view* a, const hx_view* b, int64_t p, const hx_view* result, hx_stream stream) {
    HX_API_BEGIN
    if (!b) raise(HX_ERR_INVALID_ARGUMENT, "b pointer must not be null.");
    hx::modop(0, a, b, 0, nullptr, p, result, as_stream(stream));
    HX_API_END
}
int hx_modmul_tct(const hx_view* a, int64_t b, const hx_view* p, const hx_view* result, hx_stream stream) {
    HX_API_BEGIN
    if (!p) raise(HX_ERR_INVALID_ARGUMENT, "p pointer must not be null.");
    hx::modop(0, a, nullptr, b, p, 0, result, as_stream(stream));
    HX_API_END
}
int hx_modmul_tcc(const hx_view* a, int64_t b, int64_t p, const hx_view* result, hx_stream stream) {
    HX_API_BEGIN
    hx::modop(0, a, nullptr, b, nullptr, p, result, as_stream(stream));
    HX_API_END
}
int hx_modsum_ttt(const hx_view* a, const hx_view* b, const hx_view* p, const hx_view* result, hx_stream stream) {
    HX_API_BEGIN
    if (!b) raise(HX_ERR_INVALID_ARGUMENT, "b pointer must not be null.");
    if (!p) raise(HX_ERR_INVALID_ARGUMENT, "p pointer must not be null.");
    hx::modop(1, a, b, 0, p, 0, result, as_stream(stream));
    HX_API_END
}
int hx_modsum_ttc(const hx_view* a, const hx_view* b, int64_t p, const hx_view* result, hx_stream stream) {
    HX_API_BEGIN
    if (!b) raise(HX_ERR_INVALID_ARGUMENT, "b pointer must not be null.");
    hx::modop(1, a, b, 0, nullptr, p, result, as_stream(stream));
    HX_API_END
}
int hx_modsum_tct(const hx_view* a, int64_t b, const hx_view* p, const hx_view* result, hx_stream stream) {
    HX_API_BEGIN
    if (!p) raise(HX_ERR_INVALID_ARGUMENT, "p pointer must not be null.");
    hx::modop(1, a, nullptr, b, p, 0, result, as_stream(stream));
    HX_API_END
}
int hx_modsum_tcc(const hx_view* a, int64_t b, int64_t p, const hx_view* result, hx_stream stream) {
    HX_API_BEGIN
    hx::modop(1, a, nullptr, b, nullptr, p, result, as_stream(stream));
    HX_API_END
}

int hx_mod_tt(const hx_view* a, const hx_view* b, const hx_view* result, hx_stream stream) {
    HX_API_BEGIN
    hx::modrem(a, true, 0, b, true, 0, result, as_stream(stream));
    HX_API_END
}
int hx_mod_tc(const hx_view* a, int64_t b, const hx_view* result, hx_stream stream) {
    HX_API_BEGIN
    hx::modrem(a, true, 0, nullptr, false, b, result, as_stream(stream));
    HX_API_END
}
int hx_mod_ct(int64_t a, const hx_view* b, const hx_view* result, hx_stream stream) {
    HX_API_BEGIN
    hx::modrem(nullptr, false, a, b, true, 0, result, as_stream(stream));
    HX_API_END
}

int hx_modneg_tt(const hx_view* a, const hx_view* p, const hx_view* result, hx_stream stream) {
    HX_API_BEGIN
    if (!p) raise(HX_ERR_INVALID_ARGUMENT, "p pointer must not be null.");
    hx::modop(3, a, nullptr, 0, p, 0, result, as_stream(stream));
    HX_API_END
}

int hx_modneg_tc(const hx_view* a, int64_t p, const hx_view* result, hx_stream stream) {
    HX_API_BEGIN
    hx::modop(3, a, nullptr, 0, nullptr, p, result, as_stream(stream));
    HX_API_END
}

int hx_modmul_axis_sum(const hx_view* a, const hx_view* b, const hx_view* p, const hx_view* result, int64_t axis,
                       hx_stream stream) {
    HX_API_BEGIN
    if (!a || !b || !p || !result) raise(HX_ERR_INVALID_ARGUMENT, "tensor pointers must not be null.");
    hx::modmul_axis_sum(a, b, p, result, axis, as_stream(stream));
    HX_API_END
}

int hx_rot_mac(const hx_view* acc, const hx_view* ks, const hx_view* ct, const hx_view* perm, const hx_view* diags,
               int64_t diag, const hx_view* p, hx_stream stream) {
    HX_API_BEGIN
    if (!acc || !ct || !diags || !p) raise(HX_ERR_INVALID_ARGUMENT, "tensor pointers must not be null.");
    hx::rot_mac(acc, ks, ct, perm, diags, diag, p, as_stream(stream));
    HX_API_END
}

int hx_abs(const hx_view* a, const hx_view* result, hx_stream stream) {
    HX_API_BEGIN
    if (!a || !result) raise(HX_ERR_INVALID_ARGUMENT, "tensor pointers must not be null.");
    hx::abs_values(a, result, as_stream(stream));
    HX_API_END
}

int hx_set_const(const hx_view* a, int64_t ivalue, double fvalue, hx_stream stream) {
    HX_API_BEGIN
    if (!a) raise(HX_ERR_INVALID_ARGUMENT, "tensor pointer must not be null.");
    hx::set_const(a, ivalue, fvalue, as_stream(stream));
    HX_API_END
}

int hx_axis_modsum(const hx_view* a, const hx_view* p, const hx_view* result, int64_t axis, hx_stream stream) {
    HX_API_BEGIN
    hx::axis_modsum(a, p, result, axis, as_stream(stream));
    HX_API_END
}

int hx_g_decomposition(const hx_view* a, const hx_view* result, uint64_t power, uint64_t base_bits, hx_stream stream) {
    HX_API_BEGIN
    hx::g_decomposition(a, result, power, base_bits, 0, as_stream(stream));
    HX_API_END
}

int hx_g_decomposition_tagged(const hx_view* a, const hx_view* result, uint64_t power, uint64_t base_bits,
                              uint64_t tag, hx_stream stream) {
    HX_API_BEGIN
    hx::g_decomposition(a, result, power, base_bits, tag, as_stream(stream));
    HX_API_END
}

int hx_permute(const hx_view* a, const hx_view* perms, const hx_view* result, int64_t elementwise_axis,
               int64_t perm_axis, hx_stream stream) {
    HX_API_BEGIN
    hx::permute(a, perms, result, elementwise_axis, perm_axis, false, false, as_stream(stream));
    HX_API_END
}

int hx_perms_rows_same(const hx_view* perms, int* same, hx_stream stream) {
    HX_API_BEGIN
    if (!same) raise(HX_ERR_INVALID_ARGUMENT, "same pointer must not be null.");
    *same = hx::perms_rows_same(perms, as_stream(stream)) ? 1 : 0;
    HX_API_END
}

int hx_permute_ex(const hx_view* a, const hx_view* perms, const hx_view* result, int64_t elementwise_axis,
                  int64_t perm_axis, uint32_t flags, hx_stream stream) {
    HX_API_BEGIN
    hx::permute(a, perms, result, elementwise_axis, perm_axis, (flags & HX_PERMS_TRUSTED) != 0,
                (flags & HX_PERMS_ROWS_SAME) != 0, as_stream(stream));
    HX_API_END
}

}  // extern "C"
