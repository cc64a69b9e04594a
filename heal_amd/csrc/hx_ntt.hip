// hx_ntt.hip — batched negacyclic NTT / INTT for gfx950.
//
// Semantics (reference example_impl/ntt_impl.cpp:69-170): for every column
// (i, j, t) of an [l, m, r, k] view the forward transform runs the radix-2
// Cooley–Tukey network "stage = 1, 2, 4, ..., m/2; butterfly (x[jx], x[jx+step])
// with s = tw[t, stage + u]" and then gathers out[u] = x[perm[u]]; the inverse
// scatters x[perm[u]] = a[u], runs the Gentleman–Sande network with
// s = itw[t, half + group], and multiplies by m_inv[t].  Every butterfly result
// is reduced mod q, so any schedule of the same butterflies with exact
// arithmetic gives the same output: this file changes the schedule, not the
// network.
//
// Schedule: the log2(m) stages are cut into windows of at most 8 bits (or one
// window of up to 12 bits when m <= 4096).  One launch per window ("pass")
// loads, for S sub-transforms at a time, the 2^H elements that the window's
// stages couple into LDS, runs the H stages there with one barrier per stage,
// and writes back.  Forward windows go from the top index bits down, inverse
// windows from the bottom up.  The permutation is fused into the pass that
// touches HBM last (forward) or first (inverse) as a row map.  Butterflies use
// Shoup multiplication with companions precomputed once per plan; limbs with
// q < 2^31 use 32-bit arithmetic.
//
// A plan (hx_ntt_plan) snapshots p, twiddles, perm and m_inv on the device, so
// executing it reads only the data columns and the packed twiddle table.
#include <cstring>
#include <new>
#include <vector>

#include "hx_internal.hpp"

struct LimbInfo {
    uint64_t q;
    uint64_t d;      // Möller–Granlund normalised divisor
    uint64_t v;      // and its reciprocal
    uint32_t s;
    uint32_t small;  // 1: q < 2^31, butterflies in 32-bit arithmetic
    uint64_t minv;   // m^-1 mod q (inverse plans)
    uint64_t minv_p; // its Shoup companion (32- or 64-bit)
    uint64_t pad[2];
};

struct hx_ntt_plan {
    int device;
    int inverse;
    int dtype;
    int logm;
    int64_t m;
    int64_t k;
    int perm_kind;        // PERM_* below
    void* mem;            // single device allocation
    // Intermediate buffer between passes.  A plan executes on one stream at a
    // time (like an FFT plan); the workspace grows on first use.
    mutable void* ws;
    mutable size_t ws_bytes;
    LimbInfo* limbs;      // [k]
    uint64_t* tw;         // [k][m][2]: (w mod q, companion)
    int32_t* rowmap;      // [m] (see perm_kind)
};

namespace hx {
namespace {

enum : int {
    PERM_IDENTITY = 0,
    PERM_MAP = 1,          // a permutation: rowmap[perm[u]] = u (fused)
    PERM_GATHER_FWD = 2,   // forward, not a permutation: out[u] = x[rowmap[u]] (rowmap = perm)
    PERM_SCATTER_INV = 3,  // inverse, not a permutation: x[j] = a[rowmap[j]] or old result if -1
};

constexpr int kThreads = 256;

__device__ __forceinline__ ModInfo limb_modinfo(const LimbInfo& L) {
    ModInfo mi;
    mi.M = L.q;
    mi.d = L.d;
    mi.v = L.v;
    mi.s = L.s;
    mi.pad = 0;
    return mi;
}

template <typename T>
__device__ __forceinline__ int64_t ldv(const void* base, int64_t off) {
    return (int64_t)((const T*)base)[off];
}

__device__ __forceinline__ int64_t ld_elem(const void* base, int64_t off, int elem32) {
    return elem32 ? (int64_t)((const int32_t*)base)[off] : ((const int64_t*)base)[off];
}

// Exact residue of a signed value (inputs of the NTT are non-negative; a
// negative one is reduced mathematically).
__device__ __forceinline__ uint64_t residue(const ModInfo& mi, int64_t x) {
    if (x >= 0) return reduce64(mi, (uint64_t)x);
    uint64_t r = reduce64(mi, uabs64(x));
    return r ? mi.M - r : 0;
}

// ------------------------------------------------------------ plan kernels
struct PlanSrc {
    const void* ptr;
    int64_t s0, s1;
    int32_t elem32;
};

__global__ void k_plan_limbs(PlanSrc p, PlanSrc minv, int has_minv, int64_t k, LimbInfo* limbs,
                             unsigned long long* flags) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= k) return;
    int64_t q = ld_elem(p.ptr, t * p.s0, p.elem32);
    LimbInfo L;
    memset(&L, 0, sizeof L);
    if (q <= 1) {
        atomicAdd(&flags[0], 1ull);
        q = 2;  // keeps the table well-formed; the plan is rejected on the host
    }
    ModInfo mi = make_modinfo((uint64_t)q);
    L.q = (uint64_t)q;
    L.d = mi.d;
    L.v = mi.v;
    L.s = mi.s;
    L.small = q < (int64_t(1) << 31);
    if (has_minv) {
        uint64_t mv = residue(mi, ld_elem(minv.ptr, t * minv.s0, minv.elem32));
        L.minv = mv;
        L.minv_p = L.small ? divq128(mi, 0, mv << 32) : shoup_companion(mi, mv);
    }
    limbs[t] = L;
}

__global__ void k_plan_twiddles(PlanSrc tw, int64_t k, int64_t m, const LimbInfo* limbs, uint64_t* out) {
    int64_t n = k * m;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t t = e / m, i = e - t * m;
        LimbInfo L = limbs[t];
        ModInfo mi = limb_modinfo(L);
        uint64_t w = residue(mi, ld_elem(tw.ptr, t * tw.s0 + i * tw.s1, tw.elem32));
        uint64_t wp = L.small ? divq128(mi, 0, w << 32) : shoup_companion(mi, w);
        out[2 * e] = w;
        out[2 * e + 1] = wp;
    }
}

__global__ void k_plan_perm_count(PlanSrc perm, int64_t m, int32_t* counts, unsigned long long* flags) {
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < m; u += (int64_t)gridDim.x * blockDim.x) {
        int64_t v = ld_elem(perm.ptr, u * perm.s0, perm.elem32);
        if (v < 0 || v >= m) {
            atomicAdd(&flags[1], 1ull);
            continue;
        }
        if (v != u) atomicAdd(&flags[2], 1ull);
        atomicAdd(&counts[v], 1);
    }
}

__global__ void k_plan_perm_check(const int32_t* counts, int64_t m, unsigned long long* flags) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x)
        if (counts[j] != 1) atomicAdd(&flags[3], 1ull);
}

// mode 0: map[perm[u]] = u; mode 1: map[u] = perm[u]; mode 2: map[perm[u]] = max u (map pre-filled with -1)
__global__ void k_plan_map(PlanSrc perm, int64_t m, int32_t* map, int mode) {
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < m; u += (int64_t)gridDim.x * blockDim.x) {
        int32_t v = (int32_t)ld_elem(perm.ptr, u * perm.s0, perm.elem32);
        if (mode == 0) map[v] = (int32_t)u;
        else if (mode == 1) map[u] = v;
        else atomicMax(&map[v], (int32_t)u);
    }
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t val) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = val;
}

// ------------------------------------------------------------ pass kernel
// Column addressing: a column is (i, j, t); the three axes are listed in
// `order` (innermost first) with their sizes and strides for src and dst.
// The intermediate buffer uses rows of `ncols` uint64 with the column index
// innermost.
struct ColAddr {
    const void* ptr;
    int64_t row_stride;
    int64_t cstride[3];
    int32_t kind;  // 0: int64 tensor, 1: int32 tensor, 2: uint64 intermediate
};

struct PassArgs {
    ColAddr src;
    ColAddr dst;
    ColAddr alt;             // old result, read where a scatter leaves a row untouched
    int64_t ncols;
    int64_t cdim[3];
    int32_t limb_axis;       // which of the 3 column axes is t
    int32_t inverse;
    int32_t logm;
    int32_t b0;              // lowest index bit of the window
    int32_t H;               // window bits
    int32_t stage0;          // first global stage of this pass (h for forward, g for inverse)
    int32_t S_cols, S_bot;   // tile = S_cols columns x S_bot "bot" values
    int64_t bot_count;       // 2^b0
    int64_t n_bot_tiles, n_col_tiles;
    int32_t reduce_in;
    int32_t scale_out;
    const int32_t* src_map;  // row gather on load (or null)
    const int32_t* dst_map;  // row scatter on store (or null)
    int32_t src_map_alt;     // src_map entries of -1 read `alt` at the same row
    const LimbInfo* limbs;
    const uint64_t* tw;
    int64_t m;
};

__device__ __forceinline__ int64_t col_offset(const ColAddr& c, const int64_t* coord) {
    return coord[0] * c.cstride[0] + coord[1] * c.cstride[1] + coord[2] * c.cstride[2];
}

__device__ __forceinline__ int64_t load_any(const ColAddr& c, int64_t off) {
    if (c.kind == 1) return ((const int32_t*)c.ptr)[off];
    return ((const int64_t*)c.ptr)[off];
}

__device__ __forceinline__ void store_any(const ColAddr& c, int64_t off, uint64_t v) {
    if (c.kind == 1) ((int32_t*)c.ptr)[off] = (int32_t)v;
    else ((uint64_t*)c.ptr)[off] = v;
}

__global__ __launch_bounds__(kThreads) void k_ntt_pass(PassArgs P) {
    extern __shared__ uint64_t xs[];  // [S][W + 1]
    __shared__ int64_t s_src_off[64], s_dst_off[64], s_alt_off[64], s_jbase[64];
    __shared__ int32_t s_limb[64], s_valid[64];

    const int W = 1 << P.H;
    const int WP = W + 1;
    const int S = P.S_cols * P.S_bot;

    int64_t tile = blockIdx.x;
    const int64_t col_tile = tile % P.n_col_tiles;
    tile /= P.n_col_tiles;
    const int64_t bot_tile = tile % P.n_bot_tiles;
    const int64_t top = tile / P.n_bot_tiles;

    if (threadIdx.x < S) {
        int s = threadIdx.x;
        int cs = s % P.S_cols, bs = s / P.S_cols;
        int64_t col = col_tile * P.S_cols + cs;
        int64_t bot = bot_tile * P.S_bot + bs;
        int valid = col < P.ncols && bot < P.bot_count;
        int64_t c = valid ? col : 0;
        int64_t coord[3];
        coord[0] = c % P.cdim[0];
        c /= P.cdim[0];
        coord[1] = c % P.cdim[1];
        coord[2] = c / P.cdim[1];
        int64_t lin = valid ? col : 0;
        // the intermediate buffer addresses columns by their linear index
        int64_t lcoord[3] = {lin, 0, 0};
        s_src_off[s] = P.src.kind == 2 ? col_offset(P.src, lcoord) : col_offset(P.src, coord);
        s_dst_off[s] = P.dst.kind == 2 ? col_offset(P.dst, lcoord) : col_offset(P.dst, coord);
        s_alt_off[s] = P.alt.ptr ? col_offset(P.alt, coord) : 0;
        s_limb[s] = (int32_t)coord[P.limb_axis];
        s_valid[s] = valid;
        s_jbase[s] = (top << (P.b0 + P.H)) + (valid ? bot : 0);
    }
    __syncthreads();

    // ---- load: consecutive threads take consecutive sub-problems of one row
    for (int e = threadIdx.x; e < S * W; e += blockDim.x) {
        int s = e % S, w = e / S;
        if (!s_valid[s]) continue;
        int64_t j = s_jbase[s] + ((int64_t)w << P.b0);
        int64_t x;
        if (P.src_map) {
            int32_t r = P.src_map[j];
            if (r >= 0) x = load_any(P.src, s_src_off[s] + (int64_t)r * P.src.row_stride);
            else x = load_any(P.alt, s_alt_off[s] + j * P.alt.row_stride);
        } else {
            x = load_any(P.src, s_src_off[s] + j * P.src.row_stride);
        }
        uint64_t v = (uint64_t)x;
        if (P.reduce_in) {
            LimbInfo L = P.limbs[s_limb[s]];
            if (x < 0 || (uint64_t)x >= L.q) v = residue(limb_modinfo(L), x);
        }
        xs[s * WP + w] = v;
    }
    __syncthreads();

    // ---- H radix-2 stages in LDS; a wave stays inside one sub-problem
    const int half_count = W >> 1;
    for (int st = 0; st < P.H; ++st) {
        // forward: window bit H-1-st; inverse: window bit st
        const int bit = P.inverse ? st : (P.H - 1 - st);
        const int half = 1 << bit;
        const int g = P.stage0 + st;  // global stage
        for (int bf = threadIdx.x; bf < S * half_count; bf += blockDim.x) {
            int s = bf / half_count, i = bf - s * half_count;
            if (!s_valid[s]) continue;
            int lo = ((i >> bit) << (bit + 1)) | (i & (half - 1));
            int hi = lo + half;
            int64_t jlo = s_jbase[s] + ((int64_t)lo << P.b0);
            int64_t tw_idx = P.inverse ? ((int64_t(1) << (P.logm - 1 - g)) + (jlo >> (g + 1)))
                                       : ((int64_t(1) << g) + (jlo >> (P.logm - g)));
            const int limb = s_limb[s];
            const uint64_t* twp = P.tw + 2 * ((int64_t)limb * P.m + tw_idx);
            uint64_t w = twp[0], wp = twp[1];
            const LimbInfo& L = P.limbs[limb];
            uint64_t q = L.q;
            uint64_t u = xs[s * WP + lo], v = xs[s * WP + hi];
            uint64_t nu, nv;
            if (L.small) {
                uint32_t q32 = (uint32_t)q;
                if (!P.inverse) {
                    uint32_t t = shoup_mul32((uint32_t)v, (uint32_t)w, (uint32_t)wp, q32);
                    uint32_t a = (uint32_t)u + t;
                    nu = a >= q32 ? a - q32 : a;
                    nv = (uint32_t)u >= t ? (uint32_t)u - t : (uint32_t)u + q32 - t;
                } else {
                    uint32_t a = (uint32_t)u + (uint32_t)v;
                    nu = a >= q32 ? a - q32 : a;
                    uint32_t d = (uint32_t)u >= (uint32_t)v ? (uint32_t)u - (uint32_t)v : (uint32_t)u + q32 - (uint32_t)v;
                    nv = shoup_mul32(d, (uint32_t)w, (uint32_t)wp, q32);
                }
            } else {
                if (!P.inverse) {
                    uint64_t t = shoup_mul(v, w, wp, q);
                    nu = addmod(u, t, q);
                    nv = submod(u, t, q);
                } else {
                    nu = addmod(u, v, q);
                    nv = shoup_mul(submod(u, v, q), w, wp, q);
                }
            }
            xs[s * WP + lo] = nu;
            xs[s * WP + hi] = nv;
        }
        __syncthreads();
    }

    // ---- store
    for (int e = threadIdx.x; e < S * W; e += blockDim.x) {
        int s = e % S, w = e / S;
        if (!s_valid[s]) continue;
        int64_t j = s_jbase[s] + ((int64_t)w << P.b0);
        int64_t row = P.dst_map ? (int64_t)P.dst_map[j] : j;
        uint64_t v = xs[s * WP + w];
        if (P.scale_out) {
            const LimbInfo& L = P.limbs[s_limb[s]];
            if (L.small) v = shoup_mul32((uint32_t)v, (uint32_t)L.minv, (uint32_t)L.minv_p, (uint32_t)L.q);
            else v = shoup_mul(v, L.minv, L.minv_p, L.q);
        }
        store_any(P.dst, s_dst_off[s] + row * P.dst.row_stride, v);
    }
}

// dst[u][col] = src[map[u]][col] (forward gather for non-permutation perms)
struct GatherArgs {
    ColAddr src;
    ColAddr dst;
    int64_t ncols;
    int64_t cdim[3];
    const int32_t* map;
    int64_t m;
};

__global__ void k_row_gather(GatherArgs G) {
    int64_t n = G.m * G.ncols;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t u = e / G.ncols, col = e - u * G.ncols;
        int64_t c = col, coord[3];
        coord[0] = c % G.cdim[0];
        c /= G.cdim[0];
        coord[1] = c % G.cdim[1];
        coord[2] = c / G.cdim[1];
        int64_t lc[3] = {col, 0, 0};
        int64_t so = G.src.kind == 2 ? col_offset(G.src, lc) : col_offset(G.src, coord);
        int64_t dof = G.dst.kind == 2 ? col_offset(G.dst, lc) : col_offset(G.dst, coord);
        int64_t v = load_any(G.src, so + (int64_t)G.map[u] * G.src.row_stride);
        store_any(G.dst, dof + u * G.dst.row_stride, (uint64_t)v);
    }
}

// ------------------------------------------------------------ host side
int ilog2_exact(int64_t m) {
    int l = 0;
    while ((int64_t(1) << l) < m) ++l;
    return ((int64_t(1) << l) == m) ? l : -1;
}

void require_int_view(const hx_view* v, const char* name) {
    check_view(v, name);
    if (v->dtype != HX_INT32 && v->dtype != HX_INT64)
        raise(HX_ERR_INVALID_ARGUMENT, std::string(name) + ": NTT needs int32 or int64 tensors");
}

PlanSrc plan_src(const hx_view* v) {
    PlanSrc s;
    s.ptr = v->data;
    s.s0 = v->ndim > 0 ? v->strides[0] : 0;
    s.s1 = v->ndim > 1 ? v->strides[1] : 0;
    s.elem32 = v->dtype == HX_INT32;
    return s;
}

}  // namespace

// validate_ntt_inputs (example_impl/ntt_impl.cpp:14-43), for the plan's inputs
static void validate_plan_inputs(int inverse, const hx_view* p, const hx_view* perm, const hx_view* tw,
                                 const hx_view* minv) {
    require_int_view(p, "p");
    require_int_view(perm, "perm");
    require_int_view(tw, inverse ? "inv_twiddles" : "twiddles");
    if (p->ndim != 1) raise(HX_ERR_INVALID_ARGUMENT, "Tensor 'p' must have shape [k].");
    int64_t k = p->dims[0];
    if (perm->ndim != 1) raise(HX_ERR_INVALID_ARGUMENT, "Tensor 'perm' must have shape [m].");
    int64_t m = perm->dims[0];
    if (tw->ndim != 2 || tw->dims[0] != k || tw->dims[1] != m)
        raise(HX_ERR_INVALID_ARGUMENT, "Tensor 'twiddles' must have shape [k, m].");
    if (perm->dtype != p->dtype || tw->dtype != p->dtype) raise(HX_ERR_INVALID_ARGUMENT, "NTT tensors must share one dtype");
    if (inverse) {
        if (!minv) raise(HX_ERR_INVALID_ARGUMENT, "m_inv pointer must not be null.");
        require_int_view(minv, "m_inv");
        if (minv->dtype != p->dtype) raise(HX_ERR_INVALID_ARGUMENT, "NTT tensors must share one dtype");
        if (minv->ndim != 1) raise(HX_ERR_INVALID_ARGUMENT, "Number of indices does not match tensor dimensions.");
        if (minv->dims[0] < k) raise(HX_ERR_OUT_OF_RANGE, "Index out of bounds.");
    }
    if (m > 0 && ilog2_exact(m) < 0) raise(HX_ERR_INVALID_ARGUMENT, "NTT length m must be a power of 2");
    if (m > (int64_t(1) << 30)) raise(HX_ERR_INVALID_ARGUMENT, "NTT length m too large");
}

hx_ntt_plan* ntt_plan_create(int inverse, const hx_view* p, const hx_view* perm, const hx_view* tw,
                             const hx_view* minv, hipStream_t st) {
    validate_plan_inputs(inverse, p, perm, tw, minv);
    const int64_t k = p->dims[0], m = perm->dims[0];
    hx_ntt_plan* plan = new hx_ntt_plan;
    std::memset(plan, 0, sizeof *plan);
    check_hip(hipGetDevice(&plan->device), "hipGetDevice");
    plan->inverse = inverse;
    plan->dtype = p->dtype;
    plan->m = m;
    plan->k = k;
    plan->logm = m > 0 ? ilog2_exact(m) : 0;
    size_t limb_bytes = (size_t)(k ? k : 1) * sizeof(LimbInfo);
    size_t tw_bytes = (size_t)(k ? k : 1) * (size_t)(m ? m : 1) * 16;
    size_t map_bytes = ((size_t)(m ? m : 1) * 4 + 255) / 256 * 256;
    size_t cnt_bytes = map_bytes;
    size_t flag_bytes = 256;
    size_t total = limb_bytes + tw_bytes + map_bytes + cnt_bytes + flag_bytes;
    hipError_t e = hipMalloc(&plan->mem, total);
    if (e != hipSuccess) {
        delete plan;
        check_hip(e, "hipMalloc(plan)");
    }
    char* base = (char*)plan->mem;
    plan->limbs = (LimbInfo*)base;
    plan->tw = (uint64_t*)(base + limb_bytes);
    plan->rowmap = (int32_t*)(base + limb_bytes + tw_bytes);
    int32_t* counts = (int32_t*)(base + limb_bytes + tw_bytes + map_bytes);
    unsigned long long* flags = (unsigned long long*)(base + limb_bytes + tw_bytes + map_bytes + cnt_bytes);
    try {
        check_hip(hipMemsetAsync(counts, 0, cnt_bytes + flag_bytes, st), "hipMemsetAsync(plan)");
        PlanSrc ps = plan_src(p), ms = minv ? plan_src(minv) : PlanSrc{nullptr, 0, 0, 0};
        if (k > 0) {
            hipLaunchKernelGGL(k_plan_limbs, dim3((unsigned)((k + 255) / 256)), dim3(256), 0, st, ps, ms, inverse, k,
                               plan->limbs, flags);
            if (m > 0)
                hipLaunchKernelGGL(k_plan_twiddles, dim3(grid_for(k * m, 256, 8192)), dim3(256), 0, st, plan_src(tw), k,
                                   m, plan->limbs, plan->tw);
        }
        if (m > 0) {
            unsigned g = grid_for(m, 256, 4096);
            hipLaunchKernelGGL(k_plan_perm_count, dim3(g), dim3(256), 0, st, plan_src(perm), m, counts, flags);
            hipLaunchKernelGGL(k_plan_perm_check, dim3(g), dim3(256), 0, st, counts, m, flags);
        }
        check_hip(hipGetLastError(), "plan launch");
        unsigned long long hflags[4];
        check_hip(hipMemcpyAsync(hflags, flags, sizeof hflags, hipMemcpyDeviceToHost, st), "hipMemcpyAsync(flags)");
        check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
        if (hflags[1]) raise(HX_ERR_OUT_OF_RANGE, "Index out of bounds.");
        if (hflags[0]) raise(HX_ERR_INVALID_ARGUMENT, "NTT moduli must be greater than 1");
        if (hflags[2] == 0) {
            plan->perm_kind = PERM_IDENTITY;
        } else if (hflags[3] == 0) {
            plan->perm_kind = PERM_MAP;
            hipLaunchKernelGGL(k_plan_map, dim3(grid_for(m, 256, 4096)), dim3(256), 0, st, plan_src(perm), m,
                               plan->rowmap, 0);
        } else if (!inverse) {
            plan->perm_kind = PERM_GATHER_FWD;
            hipLaunchKernelGGL(k_plan_map, dim3(grid_for(m, 256, 4096)), dim3(256), 0, st, plan_src(perm), m,
                               plan->rowmap, 1);
        } else {
            plan->perm_kind = PERM_SCATTER_INV;
            hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(m, 256, 4096)), dim3(256), 0, st, plan->rowmap, m, -1);
            hipLaunchKernelGGL(k_plan_map, dim3(grid_for(m, 256, 4096)), dim3(256), 0, st, plan_src(perm), m,
                               plan->rowmap, 2);
        }
        check_hip(hipGetLastError(), "plan map launch");
    } catch (...) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(plan->mem);
        delete plan;
        throw;
    }
    return plan;
}

void ntt_plan_destroy(hx_ntt_plan* plan) {
    if (!plan) return;
    // the plan may still be in use by queued work on any stream
    (void)hipDeviceSynchronize();
    (void)hipFree(plan->mem);
    if (plan->ws) {
        (void)hipFreeAsync(plan->ws, nullptr);
        (void)hipDeviceSynchronize();
    }
    delete plan;
}

namespace {

ColAddr tensor_addr(const hx_view* v, const int* order) {
    ColAddr c;
    c.ptr = v->data;
    c.row_stride = v->strides[1];
    const int axes[3] = {0, 2, 3};
    for (int i = 0; i < 3; ++i) c.cstride[i] = v->strides[axes[order[i]]];
    c.kind = v->dtype == HX_INT32 ? 1 : 0;
    return c;
}

ColAddr tmp_addr(void* ptr, int64_t ncols) {
    ColAddr c;
    c.ptr = ptr;
    c.row_stride = ncols;
    c.cstride[0] = 1;
    c.cstride[1] = 0;
    c.cstride[2] = 0;
    c.kind = 2;
    return c;
}

void launch_pass(PassArgs& P, hipStream_t st) {
    const int W = 1 << P.H;
    int S = P.S_cols * P.S_bot;
    int64_t ntop = int64_t(1) << (P.logm - P.b0 - P.H);
    P.n_bot_tiles = (P.bot_count + P.S_bot - 1) / P.S_bot;
    P.n_col_tiles = (P.ncols + P.S_cols - 1) / P.S_cols;
    int64_t blocks = ntop * P.n_bot_tiles * P.n_col_tiles;
    if (blocks > 0x7fffffff) raise(HX_ERR_INVALID_ARGUMENT, "NTT too large for one launch");
    size_t lds = (size_t)S * (W + 1) * 8;
    hipLaunchKernelGGL(k_ntt_pass, dim3((unsigned)blocks), dim3(kThreads), lds, st, P);
    check_hip(hipGetLastError(), "ntt pass launch");
}

}  // namespace

// Executes a plan on a (and writes result); both [l, m, r, k] views.
void ntt_plan_exec(const hx_ntt_plan* plan, const hx_view* a, const hx_view* result, hipStream_t st) {
    if (!plan) raise(HX_ERR_INVALID_ARGUMENT, "plan pointer must not be null.");
    require_int_view(a, "a");
    require_int_view(result, "result");
    if (a->ndim != 4) raise(HX_ERR_INVALID_ARGUMENT, "Input tensor 'a' must have shape [l, m, r, k].");
    if (!dims_equal(a, result)) raise(HX_ERR_INVALID_ARGUMENT, "Output tensor must have the same shape as input tensor.");
    if (a->dims[3] != plan->k) raise(HX_ERR_INVALID_ARGUMENT, "Tensor 'p' must have shape [k].");
    if (a->dims[1] != plan->m) raise(HX_ERR_INVALID_ARGUMENT, "Tensor 'perm' must have shape [m].");
    if (a->dtype != plan->dtype || result->dtype != plan->dtype)
        raise(HX_ERR_INVALID_ARGUMENT, "NTT tensors must share one dtype");
    int dev = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (dev != plan->device) raise(HX_ERR_INVALID_ARGUMENT, "plan belongs to another device");
    const int64_t m = plan->m;
    const int64_t ncols = a->dims[0] * a->dims[2] * a->dims[3];
    if (ncols == 0 || m == 0) return;

    // column axes (l, r, k) ordered by |stride| in a, innermost first, so that
    // consecutive columns of a tile are close in memory
    int order[3] = {2, 1, 0};  // indices into (l, r, k)
    {
        const int64_t sa[3] = {a->strides[0], a->strides[2], a->strides[3]};
        const int64_t dm[3] = {a->dims[0], a->dims[2], a->dims[3]};
        auto key = [&](int i) { return dm[i] == 1 ? INT64_MAX : (sa[i] < 0 ? -sa[i] : sa[i]); };
        for (int i = 0; i < 3; ++i)
            for (int j = i + 1; j < 3; ++j)
                if (key(order[j]) < key(order[i])) std::swap(order[i], order[j]);
    }
    const int64_t dims3[3] = {a->dims[0], a->dims[2], a->dims[3]};
    PassArgs base;
    std::memset(&base, 0, sizeof base);
    base.ncols = ncols;
    for (int i = 0; i < 3; ++i) {
        base.cdim[i] = dims3[order[i]];
        if (order[i] == 2) base.limb_axis = i;
    }
    base.inverse = plan->inverse;
    base.logm = plan->logm;
    base.limbs = plan->limbs;
    base.tw = plan->tw;
    base.m = m;
    const ColAddr A = tensor_addr(a, order), R = tensor_addr(result, order);
    // are columns adjacent in memory (k-last layouts) or rows (limb-major)?
    const int64_t inner_col_stride = A.cstride[0] < 0 ? -A.cstride[0] : A.cstride[0];
    const bool rows_contiguous = (a->strides[1] == 1) && inner_col_stride != 1;

    // windows: one pass up to 12 bits, else windows of <= 8 bits
    const int logm = plan->logm;
    std::vector<int> win;
    if (logm <= 12) {
        win.push_back(logm);
    } else {
        int npass = (logm + 7) / 8;
        for (int i = 0; i < npass; ++i) win.push_back(logm / npass + (i < logm % npass ? 1 : 0));
    }
    const int npass = (int)win.size();

    void* tmp = nullptr;
    const bool need_tmp = npass > 1 || plan->perm_kind == PERM_GATHER_FWD || plan->perm_kind == PERM_SCATTER_INV;
    if (need_tmp) {
        // the plan's workspace, grown outside graph capture and reused afterwards
        size_t bytes = (size_t)m * (size_t)ncols * 8 * (plan->perm_kind >= PERM_GATHER_FWD ? 2 : 1);
        if (plan->ws_bytes < bytes) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            check_hip(hipStreamIsCapturing(st, &cs), "hipStreamIsCapturing");
            if (cs != hipStreamCaptureStatusNone)
                raise(HX_ERR_RUNTIME, "NTT plan workspace must be sized by one execution before graph capture");
            if (plan->ws) check_hip(hipFreeAsync(plan->ws, st), "hipFreeAsync(ntt workspace)");
            plan->ws = nullptr;
            plan->ws_bytes = 0;
            check_hip(hipMallocAsync(&plan->ws, bytes, st), "hipMallocAsync(ntt workspace)");
            plan->ws_bytes = bytes;
        }
        tmp = plan->ws;
    }
    void* tmp2 = (need_tmp && plan->perm_kind >= PERM_GATHER_FWD) ? (char*)tmp + (size_t)m * ncols * 8 : nullptr;
    const ColAddr T1 = need_tmp ? tmp_addr(tmp, ncols) : ColAddr{};

    // inverse with a non-permutation perm: stage the scattered input first
    ColAddr in_src = A;
    if (plan->perm_kind == PERM_SCATTER_INV) {
        // x[j] = a[map[j]] or, for rows no perm entry hits, the old result row.
        // A pass with an empty window (H = 0, b0 = 0) visits every row j once
        // as its own "top" index and runs no stages.
        PassArgs G = base;
        G.src = A;
        G.alt = R;
        G.dst = tmp_addr(tmp2, ncols);
        G.src_map = plan->rowmap;
        G.src_map_alt = 1;
        G.H = 0;
        G.b0 = 0;
        G.bot_count = 1;
        G.stage0 = 0;
        G.S_cols = 64;
        G.S_bot = 1;
        G.reduce_in = 0;
        G.scale_out = 0;
        launch_pass(G, st);
        in_src = tmp_addr(tmp2, ncols);
    }

    int done_bits = 0;
    for (int ip = 0; ip < npass; ++ip) {
        PassArgs P = base;
        const int H = win[ip];
        P.H = H;
        if (!plan->inverse) {
            P.stage0 = done_bits;            // forward: top bits first
            P.b0 = logm - done_bits - H;
        } else {
            P.stage0 = done_bits;            // inverse: bottom bits first
            P.b0 = done_bits;
        }
        P.bot_count = int64_t(1) << P.b0;
        const bool first = ip == 0, last = ip == npass - 1;
        P.src = first ? in_src : T1;
        P.dst = last ? R : T1;
        if (last && !plan->inverse && plan->perm_kind == PERM_GATHER_FWD) P.dst = tmp_addr(tmp2, ncols);
        P.reduce_in = first && (plan->inverse || m > 1);
        P.scale_out = last && plan->inverse;
        if (first && plan->inverse && plan->perm_kind == PERM_MAP) P.src_map = plan->rowmap;
        if (last && !plan->inverse && plan->perm_kind == PERM_MAP) P.dst_map = plan->rowmap;
        // tile shape
        const int W = 1 << H;
        int S = W >= 256 ? 8 : std::max(1, std::min(64, 2048 / W));
        if (npass == 1 && W > 256) S = std::max(1, 4096 / W);
        if (rows_contiguous && P.bot_count >= S) {
            P.S_cols = 1;
            P.S_bot = S;
        } else {
            P.S_cols = S;
            P.S_bot = 1;
        }
        launch_pass(P, st);
        done_bits += H;
    }

    if (!plan->inverse && plan->perm_kind == PERM_GATHER_FWD) {
        GatherArgs G;
        G.src = tmp_addr(tmp2, ncols);
        G.dst = R;
        G.ncols = ncols;
        for (int i = 0; i < 3; ++i) G.cdim[i] = base.cdim[i];
        G.map = plan->rowmap;
        G.m = m;
        hipLaunchKernelGGL(k_row_gather, dim3(grid_for(m * ncols, 256, 8192)), dim3(256), 0, st, G);
        check_hip(hipGetLastError(), "row gather launch");
    }
}

}  // namespace hx

using namespace hx;

extern "C" {

int hx_ntt_plan_create(hx_ntt_plan** out, int inverse, const hx_view* p, const hx_view* perm, const hx_view* twiddles,
                       const hx_view* m_inv, hx_stream stream) {
    HX_API_BEGIN
    if (!out) raise(HX_ERR_INVALID_ARGUMENT, "out pointer must not be null.");
    *out = nullptr;
    if (!p) raise(HX_ERR_INVALID_ARGUMENT, "p pointer must not be null.");
    if (!perm) raise(HX_ERR_INVALID_ARGUMENT, "perm pointer must not be null.");
    if (!twiddles) raise(HX_ERR_INVALID_ARGUMENT, "twiddles pointer must not be null.");
    *out = ntt_plan_create(inverse ? 1 : 0, p, perm, twiddles, m_inv, as_stream(stream));
    HX_API_END
}

int hx_ntt_plan_exec(const hx_ntt_plan* plan, const hx_view* a, const hx_view* result, hx_stream stream) {
    HX_API_BEGIN
    ntt_plan_exec(plan, a, result, as_stream(stream));
    HX_API_END
}

int hx_ntt_plan_destroy(hx_ntt_plan* plan) {
    HX_API_BEGIN
    ntt_plan_destroy(plan);
    HX_API_END
}

static int ntt_once(int inverse, const hx_view* a, const hx_view* p, const hx_view* perm, const hx_view* tw,
                    const hx_view* minv, const hx_view* result, hx_stream stream) {
    HX_API_BEGIN
    // the reference validates the data tensors before the tables (ntt_impl.cpp:24-42)
    if (!a) raise(HX_ERR_INVALID_ARGUMENT, "a pointer must not be null.");
    if (!result) raise(HX_ERR_INVALID_ARGUMENT, "result pointer must not be null.");
    if (!p || !perm || !tw) raise(HX_ERR_INVALID_ARGUMENT, "NTT table pointer must not be null.");
    check_view(a, "a");
    check_view(result, "result");
    if (a->ndim != 4) raise(HX_ERR_INVALID_ARGUMENT, "Input tensor 'a' must have shape [l, m, r, k].");
    if (!dims_equal(a, result)) raise(HX_ERR_INVALID_ARGUMENT, "Output tensor must have the same shape as input tensor.");
    check_view(p, "p");
    if (p->ndim != 1 || p->dims[0] != a->dims[3]) raise(HX_ERR_INVALID_ARGUMENT, "Tensor 'p' must have shape [k].");
    check_view(perm, "perm");
    if (perm->ndim != 1 || perm->dims[0] != a->dims[1]) raise(HX_ERR_INVALID_ARGUMENT, "Tensor 'perm' must have shape [m].");
    hipStream_t st = as_stream(stream);
    hx_ntt_plan* plan = ntt_plan_create(inverse, p, perm, tw, minv, st);
    try {
        ntt_plan_exec(plan, a, result, st);
    } catch (...) {
        ntt_plan_destroy(plan);
        throw;
    }
    // the plan's memory must outlive the queued passes
    check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    ntt_plan_destroy(plan);
    HX_API_END
}

int hx_ntt(const hx_view* a, const hx_view* p, const hx_view* perm, const hx_view* twiddles, const hx_view* log2p_list,
           const hx_view* mu_list, const hx_view* result, hx_stream stream) {
    (void)log2p_list;
    (void)mu_list;
    return ntt_once(0, a, p, perm, twiddles, nullptr, result, stream);
}

int hx_intt(const hx_view* a, const hx_view* p, const hx_view* perm, const hx_view* inv_twiddles, const hx_view* m_inv,
            const hx_view* log2p_list, const hx_view* mu_list, const hx_view* result, hx_stream stream) {
    (void)log2p_list;
    (void)mu_list;
    return ntt_once(1, a, p, perm, inv_twiddles, m_inv, result, stream);
}

}  // extern "C"
