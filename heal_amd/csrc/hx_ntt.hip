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
// window of up to 11 bits when m <= 2048).  One launch per window ("pass")
// handles, for S sub-transforms at a time, the 2^H elements that the window's
// stages couple: windows of 6..8 bits run in registers, 16 values per lane,
// with one LDS exchange (k_ntt_pass_r16); other windows run in LDS with one
// barrier per stage (k_ntt_pass).  Forward windows go from the top index bits
// down, inverse windows from the bottom up.  The permutation is fused into the pass that
// touches HBM last (forward) or first (inverse) as a row map.  Butterflies use
// Shoup multiplication with companions precomputed once per plan; limbs with
// q < 2^31 use 32-bit arithmetic.
//
// A plan (hx_ntt_plan) snapshots p, twiddles, perm and m_inv on the device, so
// executing it reads only the data columns and the packed twiddle table.
#include <new>
#include <vector>

#include "hx_ntt_kernels.hpp"

namespace hx {


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
                                   m, plan->limbs, inverse, plan->tw);
        }
        if (m > 0) {
            unsigned g = grid_for(m, 256, 4096);
            hipLaunchKernelGGL(k_plan_perm_count, dim3(g), dim3(256), 0, st, plan_src(perm), m, counts, flags);
            hipLaunchKernelGGL(k_plan_perm_check, dim3(g), dim3(256), 0, st, counts, m, flags);
            if (plan->logm <= 30)
                hipLaunchKernelGGL(k_plan_bitrev, dim3(g), dim3(256), 0, st, plan_src(perm), m, plan->logm, flags);
        }
        check_hip(hipGetLastError(), "plan launch");
        unsigned long long hflags[5];
        check_hip(hipMemcpyAsync(hflags, flags, sizeof hflags, hipMemcpyDeviceToHost, st), "hipMemcpyAsync(flags)");
        check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
        if (hflags[1]) raise(HX_ERR_OUT_OF_RANGE, "Index out of bounds.");
        if (hflags[0]) raise(HX_ERR_INVALID_ARGUMENT, "NTT moduli must be greater than 1");
        if (hflags[2] == 0) {
            plan->perm_kind = PERM_IDENTITY;
        } else if (hflags[3] == 0) {
            plan->perm_kind = PERM_MAP;
            plan->perm_bitrev = m > 1 && plan->logm <= 30 && hflags[4] == 0;
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

    // windows: one pass up to 11 bits, else near-equal windows of <= 8 bits
    // (12..16 bits: two register-radix-16 passes)
    const int logm = plan->logm;
    std::vector<int> win;
    if (logm <= 11) {
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
    // the intermediate follows the data: columns innermost for limb-last
    // layouts, rows innermost for limb-major ones (coalesced either way)
    auto tmp_at = [&](void* ptr) { return rows_contiguous ? tmp_addr_rows(ptr, m) : tmp_addr(ptr, ncols); };
    const ColAddr T1 = need_tmp ? tmp_at(tmp) : ColAddr{};

    // inverse with a non-permutation perm: stage the scattered input first
    ColAddr in_src = A;
    if (plan->perm_kind == PERM_SCATTER_INV) {
        // x[j] = a[map[j]] or, for rows no perm entry hits, the old result row.
        // A pass with an empty window (H = 0, b0 = 0) visits every row j once
        // as its own "top" index and runs no stages.
        PassArgs G = base;
        G.src = A;
        G.alt = R;
        G.dst = tmp_at(tmp2);
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
        in_src = tmp_at(tmp2);
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
        if (last && !plan->inverse && plan->perm_kind == PERM_GATHER_FWD) P.dst = tmp_at(tmp2);
        P.reduce_in = first && (plan->inverse || m > 1);
        P.scale_out = last && plan->inverse;
        if (first && plan->inverse && plan->perm_kind == PERM_MAP) P.src_map = plan->rowmap;
        if (last && !plan->inverse && plan->perm_kind == PERM_MAP) P.dst_map = plan->rowmap;
        // radix-16 passes hand lazy residues to a following radix-16 pass
        P.lazy_out = !last && r16_window(H) && r16_window(win[ip + 1]);
        const bool col_tiles = !(rows_contiguous && P.bot_count >= 8);
        configure_tile(P, col_tiles, npass);
        // a tile of whole columns whose rows are contiguous: lanes of a sub-problem
        // take adjacent rows (coalesced), not adjacent columns
        P.lane_major = col_tiles && rows_contiguous;
        // the bit-reversal row map of rows-contiguous columns, fused into the
        // pass over the low bits (the forward's last, the inverse's first): tiles
        // of tops whose permuted rows are adjacent
        const bool map_side = plan->inverse ? (P.src_map != nullptr && P.src.row_stride == 1)
                                            : (P.dst_map != nullptr && P.dst.row_stride == 1);
        P.top_bitrev = plan->perm_bitrev && rows_contiguous && map_side && P.b0 == 0 && r16_window(H) &&
                       logm - H >= 5 && npass > 1 && P.src.row_stride == 1 && P.dst.row_stride == 1;
        // other passes with the bit-reversal map compute its rows instead of loading them
        P.map_bitrev = plan->perm_bitrev && (P.src_map || P.dst_map);
        launch_pass(P, st);
        done_bits += H;
    }

    if (!plan->inverse && plan->perm_kind == PERM_GATHER_FWD) {
        GatherArgs G;
        G.src = tmp_at(tmp2);
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
