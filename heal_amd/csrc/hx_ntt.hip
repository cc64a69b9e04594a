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
#include <mutex>
#include <new>
#include <string>
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
                             const hx_view* minv, hipStream_t caller) {
    validate_plan_inputs(inverse, p, perm, tw, minv);
    const int64_t k = p->dims[0], m = perm->dims[0];
    // preparation synchronises once; inside a graph capture it runs on the side stream
    SyncScope sc(caller);
    hipStream_t st = sc.stream();
    hx_ntt_plan* plan = new hx_ntt_plan;
    std::memset(plan, 0, sizeof *plan);
    check_hip(hipGetDevice(&plan->device), "hipGetDevice");
    plan->inverse = inverse;
    plan->dtype = p->dtype;
    plan->m = m;
    plan->k = k;
    plan->logm = m > 0 ? ilog2_exact(m) : 0;
    size_t limb_bytes = ((size_t)(k ? k : 1) * sizeof(LimbInfo) + 255) / 256 * 256;
    size_t tw_bytes = (size_t)(k ? k : 1) * (size_t)(m ? m : 1) * 16;
    size_t map_bytes = ((size_t)(m ? m : 1) * 4 + 255) / 256 * 256;
    size_t cnt_bytes = map_bytes;
    size_t flag_bytes = 256;
    size_t total = limb_bytes + tw_bytes + map_bytes + cnt_bytes + flag_bytes;
    check_hip(hipMallocAsync(&plan->mem, total, st), "hipMallocAsync(plan)");
    char* base = (char*)plan->mem;
    plan->limbs = (LimbInfo*)base;
    plan->tw = (uint64_t*)(base + limb_bytes);
    plan->rowmap = (int32_t*)(base + limb_bytes + tw_bytes);
    int32_t* counts = (int32_t*)(base + limb_bytes + tw_bytes + map_bytes);
    unsigned long long* flags = (unsigned long long*)(base + limb_bytes + tw_bytes + map_bytes + cnt_bytes);
    // moduli whose reference sums wrap for reduced values: int64 sums of two
    // residues exceed 2^63 for q > 2^62; the int32 inverse's u + q - v exceeds
    // 2^31 for q > 2^30 (the int32 forward sums in int64)
    const int64_t qmax = p->dtype == HX_INT64 ? (int64_t(1) << 62) : (inverse ? (int64_t(1) << 30) : INT64_MAX);
    try {
        fill_zero(counts, cnt_bytes + flag_bytes, st);
        PlanSrc ps = plan_src(p), ms = minv ? plan_src(minv) : PlanSrc{nullptr, 0, 0, 0};
        if (k > 0) {
            hipLaunchKernelGGL(k_plan_limbs, dim3((unsigned)((k + 255) / 256)), dim3(256), 0, st, ps, ms, inverse, k,
                               qmax, plan->limbs, flags);
            if (m > 0)
                hipLaunchKernelGGL(k_plan_twiddles, dim3(grid_for(k * m, 256, 8192)), dim3(256), 0, st, plan_src(tw), k,
                                   m, plan->limbs, inverse, plan->tw, flags);
        }
        if (m > 0) {
            unsigned g = grid_for(m, 256, 4096);
            hipLaunchKernelGGL(k_plan_perm_count, dim3(g), dim3(256), 0, st, plan_src(perm), m, counts, flags);
            hipLaunchKernelGGL(k_plan_perm_check, dim3(g), dim3(256), 0, st, counts, m, flags);
            if (plan->logm <= 30)
                hipLaunchKernelGGL(k_plan_bitrev, dim3(g), dim3(256), 0, st, plan_src(perm), m, plan->logm, flags);
        }
        check_hip(hipGetLastError(), "plan launch");
        unsigned long long hflags[7];
        check_hip(hipMemcpyAsync(hflags, flags, sizeof hflags, hipMemcpyDeviceToHost, st), "hipMemcpyAsync(flags)");
        sc.sync();
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
        plan->ref_all = (hflags[5] || hflags[6]) && k > 0 && m > 0;
        if (plan->ref_all) {
            // the reference's network reads the tables as given
            const size_t rb = (size_t)k * m * 8 + (size_t)k * 8;
            void* raw = nullptr;
            check_hip(hipMallocAsync(&raw, rb, st), "hipMallocAsync(plan tables)");
            plan->raw_tw = (int64_t*)raw;
            plan->raw_minv = inverse ? plan->raw_tw + k * m : nullptr;
            hipLaunchKernelGGL(k_plan_raw, dim3(grid_for(k * m, 256, 8192)), dim3(256), 0, st, plan_src(tw), ms,
                               inverse, k, m, plan->raw_tw, plan->raw_minv);
            check_hip(hipGetLastError(), "plan table launch");
        }
        plan->wss = new Workspaces;
        plan->wss->device = plan->device;
        // the plan is complete before any stream can use it
        sc.sync();
    } catch (...) {
        (void)hipStreamSynchronize(st);
        if (plan->raw_tw) (void)hipFreeAsync(plan->raw_tw, st);
        (void)hipFreeAsync(plan->mem, st);
        (void)hipStreamSynchronize(st);
        delete plan->wss;
        delete plan;
        throw;
    }
    return plan;
}

// The plan's memory is freed after the work already queued with it on any
// stream (stream-ordered, no host wait).  Graphs captured with the plan must
// be destroyed first.
void ntt_plan_destroy(hx_ntt_plan* plan) {
    if (!plan) return;
    std::vector<hipEvent_t> events;
    std::vector<void*> ptrs{plan->mem, plan->raw_tw};
    if (plan->wss) {
        std::lock_guard<std::mutex> lk(plan->wss->mu);
        plan->wss->release(events, ptrs);
    }
    release_after(plan->device, events, ptrs);
    delete plan->wss;
    delete plan;
}

// Workspace of one execution: [flags: ncols uint32][T1][T2][fixup scratch]
//   T1   the intermediate between passes, also the fixup kernel's scratch;
//   T2   staging of a non-permutation perm (forward gather, inverse scatter);
//   fixup scratch: kRefBlocks rows of m int64 when there is no T1.
constexpr int64_t kRefBlocks = 1024;

struct ExecShape {
    int64_t ncols;
    int win[4];  // log2 m <= 30: at most four windows
    int npass;
    bool need_tmp;
    size_t tmp_bytes;   // T1 (+ T2)
    size_t data_bytes;  // tmp_bytes + fixup scratch
    size_t flag_bytes;
    int64_t ref_blocks;
};

static ExecShape exec_shape(const hx_ntt_plan* plan, const hx_view* a) {
    ExecShape E;
    E.ncols = a->dims[0] * a->dims[2] * a->dims[3];
    const int64_t m = plan->m;
    const int logm = plan->logm;
    if (logm <= 11) {
        E.npass = 1;
        E.win[0] = logm;
    } else {
        E.npass = (logm + 7) / 8;
        for (int i = 0; i < E.npass; ++i) E.win[i] = logm / E.npass + (i < logm % E.npass ? 1 : 0);
    }
    const bool two = plan->perm_kind >= PERM_GATHER_FWD;
    E.need_tmp = E.npass > 1 || two;
    E.tmp_bytes = E.need_tmp ? (size_t)m * (size_t)E.ncols * 8 * (two ? 2 : 1) : 0;
    E.ref_blocks = std::min<int64_t>(E.ncols, kRefBlocks);
    E.data_bytes = E.tmp_bytes + (E.need_tmp ? 0 : (size_t)E.ref_blocks * (size_t)m * 8);
    E.flag_bytes = ((size_t)E.ncols * 4 + 255) / 256 * 256;
    return E;
}

static void validate_exec(const hx_ntt_plan* plan, const hx_view* a, const hx_view* result) {
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
}

// Executes a plan on a (and writes result); both [l, m, r, k] views.  flags
// (E.flag_bytes) must be zero on entry and are zero again when the queued work
// has run; data holds E.data_bytes.
static void ntt_plan_run(const hx_ntt_plan* plan, const hx_view* a, const hx_view* result, const ExecShape& E,
                         uint32_t* flags, void* data, hipStream_t st) {
    const int64_t m = plan->m;
    const int64_t ncols = E.ncols;
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

    const int* win = E.win;
    const int npass = E.npass;
    const int logm = plan->logm;
    void* tmp = E.need_tmp ? data : nullptr;
    void* tmp2 = (E.need_tmp && plan->perm_kind >= PERM_GATHER_FWD) ? (char*)tmp + (size_t)m * ncols * 8 : nullptr;
    // the intermediate follows the data: columns innermost for limb-last
    // layouts, rows innermost for limb-major ones (coalesced either way)
    auto tmp_at = [&](void* ptr) { return rows_contiguous ? tmp_addr_rows(ptr, m) : tmp_addr(ptr, ncols); };
    const ColAddr T1 = E.need_tmp ? tmp_at(tmp) : ColAddr{};

    if (plan->ref_all) {
        fill_u32(flags, ncols, 1u, st);
    } else {
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
            // the first pass flags columns whose reference result is not the
            // residue; the last one leaves them to k_ntt_ref (a gather after it
            // does the same)
            P.colflag = flags;
            P.flag_set = first;
            P.flag_skip = last && !(!plan->inverse && plan->perm_kind == PERM_GATHER_FWD);
            P.flag_kind = plan->inverse ? FLAG_INV : (plan->dtype == HX_INT32 ? FLAG_FWD32 : FLAG_FWD64);
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
            G.colflag = flags;
            hipLaunchKernelGGL(k_row_gather, dim3(grid_for(m * ncols, 256, 8192)), dim3(256), 0, st, G);
            check_hip(hipGetLastError(), "row gather launch");
        }
    }

    // the reference's own network on the flagged columns (most launches find none)
    RefArgs F;
    std::memset(&F, 0, sizeof F);
    F.src = A;
    F.old = plan->perm_kind == PERM_SCATTER_INV ? R : ColAddr{};
    F.dst = R;
    F.scratch = (int64_t*)(E.need_tmp ? tmp : data);
    F.colflag = flags;
    F.rowmap = plan->rowmap;
    F.perm_kind = plan->perm_kind;
    F.inverse = plan->inverse;
    F.is32 = plan->dtype == HX_INT32;
    F.logm = logm;
    F.m = m;
    F.ncols = ncols;
    for (int i = 0; i < 3; ++i) F.cdim[i] = base.cdim[i];
    F.limb_axis = base.limb_axis;
    F.limbs = plan->limbs;
    F.tw = plan->tw;
    F.raw_tw = plan->raw_tw;
    F.raw_minv = plan->raw_minv;
    hipLaunchKernelGGL(k_ntt_ref, dim3((unsigned)E.ref_blocks), dim3(256), 0, st, F);
    check_hip(hipGetLastError(), "reference-semantics launch");
}

// Executes with the plan's own workspace for st (see hx::Workspaces).
void ntt_plan_exec(const hx_ntt_plan* plan, const hx_view* a, const hx_view* result, hipStream_t st) {
    validate_exec(plan, a, result);
    const ExecShape E = exec_shape(plan, a);
    if (E.ncols == 0 || plan->m == 0) return;
    std::lock_guard<std::mutex> lk(plan->wss->mu);
    Workspaces::Ws w = plan->wss->acquire(st, E.flag_bytes, E.data_bytes);
    ntt_plan_run(plan, a, result, E, (uint32_t*)w.flags, w.data, st);
    plan->wss->note(st);
}

size_t ntt_workspace_bytes(const hx_ntt_plan* plan, const hx_view* a, const hx_view* result) {
    validate_exec(plan, a, result);
    const ExecShape E = exec_shape(plan, a);
    if (E.ncols == 0 || plan->m == 0) return 0;
    return E.flag_bytes + E.data_bytes;
}

// Executes with a caller's workspace of ntt_workspace_bytes(): flags are
// cleared first, so its contents on entry do not matter.
void ntt_plan_exec_ws(const hx_ntt_plan* plan, const hx_view* a, const hx_view* result, void* ws, size_t bytes,
                      hipStream_t st) {
    validate_exec(plan, a, result);
    const ExecShape E = exec_shape(plan, a);
    if (E.ncols == 0 || plan->m == 0) return;
    if (!ws) raise(HX_ERR_INVALID_ARGUMENT, "workspace pointer must not be null.");
    if ((uintptr_t)ws % 256) raise(HX_ERR_INVALID_ARGUMENT, "workspace must be 256-byte aligned");
    if (bytes < E.flag_bytes + E.data_bytes) raise(HX_ERR_INVALID_ARGUMENT, "workspace too small for this execution");
    fill_zero(ws, E.flag_bytes, st);
    ntt_plan_run(plan, a, result, E, (uint32_t*)ws, (char*)ws + E.flag_bytes, st);
}

// ------------------------------------------------------ no-allocation execution
// The workspace of hx_ntt_plan_exec_noalloc: a device array of fixed size,
// one copy per GPU, held until the process ends.  Its host-side name is not a
// device address: kernels and fills get the address hipGetSymbolAddress
// returns on each device, taken once per device on the first call there.
static __device__ __align__(256) unsigned char scratch[HX_NTT_NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static std::vector<void*> g_scratch_ptr;  // per device

// The current device's copy of `scratch` (callers have checked that the
// current device is the plan's).
static void* scratch_ptr(int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!ptr) raise(HX_ERR_RUNTIME, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

// Executes a plan with `scratch` as its workspace: no device memory is
// allocated.  The flags are cleared by a fill kernel first; the rest of the
// scratch needs no initial contents.  Every call on a device shares that
// device's scratch, so calls that overlap in time on one GPU race.
void ntt_plan_exec_noalloc(const hx_ntt_plan* plan, const hx_view* a, const hx_view* result, hipStream_t st) {
    validate_exec(plan, a, result);
    const ExecShape E = exec_shape(plan, a);
    if (E.ncols == 0 || plan->m == 0) return;
    const size_t need = E.flag_bytes + E.data_bytes;
    if (need > (size_t)HX_NTT_NOALLOC_SCRATCH_BYTES)
        raise(HX_ERR_INVALID_ARGUMENT, "this transform needs " + std::to_string(need) +
                                           " bytes of workspace; hx_ntt_plan_exec_noalloc holds " +
                                           std::to_string((size_t)HX_NTT_NOALLOC_SCRATCH_BYTES));
    char* ptr = (char*)scratch_ptr(plan->device);
    fill_zero(ptr, E.flag_bytes, st);
    ntt_plan_run(plan, a, result, E, (uint32_t*)ptr, ptr + E.flag_bytes, st);
}

// ------------------------------------------------------------ VALU ceiling
// Throughput of the passes' 60-bit butterfly (MODE_L64, forward) with its
// operands in registers: every thread runs 16 values as 8 independent
// butterflies per stage, each with its own twiddle pair, two stages per round
// (one that reduces its upper inputs and one that does not, as the passes
// alternate them), at 8 waves per SIMD.  No memory traffic but one store per
// thread at the end, so its rate is the arithmetic ceiling of a pass of these
// butterflies on this GPU.
constexpr int kRateRounds = 512;

__global__ __launch_bounds__(256) void k_butterfly_rate(uint64_t* out, uint64_t q, uint64_t seed) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) x[i] = ((t + 1) * 0x9e3779b97f4a7c15ull + (uint64_t)i * seed) % q;
    uint64_t w[8], wp[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        w[i] = ((seed ^ t) + (uint64_t)i * 0x1234567ull) % q;
        wp[i] = (w[i] * 0x2545f4914f6cdd1dull) | 1u;
    }
    for (int r = 0; r < kRateRounds; ++r) {
#pragma unroll
        for (int i = 0; i < 16; i += 2) butterfly<MODE_L64, false, true>(x[i], x[i + 1], w[i / 2], wp[i / 2], q);
#pragma unroll
        for (int i = 0; i < 16; i += 2) butterfly<MODE_L64, false, false>(x[i], x[i + 1], w[i / 2], wp[i / 2], q);
    }
    uint64_t h = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) h ^= x[i];
    out[t] = h;
}

// Butterflies per second of k_butterfly_rate over the whole GPU (best of
// three launches of 8 workgroups of 4 waves per CU).
double butterfly_rate(hipStream_t st) {
    int dev = 0, cus = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    const int blocks = std::max(1, cus) * 8, threads = 256;
    uint64_t* out = nullptr;
    check_hip(hipMallocAsync((void**)&out, (size_t)blocks * threads * 8, st), "hipMallocAsync(rate)");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float best = 1e30f;
    try {
        check_hip(hipEventCreate(&e0), "hipEventCreate");
        check_hip(hipEventCreate(&e1), "hipEventCreate");
        const uint64_t q = 1152921504606844417ull;  // 2^60 - 2^15 + 1
        for (int r = 0; r < 4; ++r) {
            check_hip(hipEventRecord(e0, st), "hipEventRecord");
            hipLaunchKernelGGL(k_butterfly_rate, dim3(blocks), dim3(threads), 0, st, out, q, (uint64_t)(r + 3));
            check_hip(hipGetLastError(), "butterfly rate launch");
            check_hip(hipEventRecord(e1, st), "hipEventRecord");
            check_hip(hipEventSynchronize(e1), "hipEventSynchronize");
            float ms = 0;
            check_hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
            if (r > 0) best = std::min(best, ms);  // the first launch warms up
        }
    } catch (...) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        (void)hipFreeAsync(out, st);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    check_hip(hipFreeAsync(out, st), "hipFreeAsync(rate)");
    const double bflies = (double)blocks * threads * kRateRounds * 16;
    return bflies / (best * 1e-3);
}

}  // namespace hx

using namespace hx;

extern "C" {

int hx_butterfly_rate(double* butterflies_per_s, hx_stream stream) {
    HX_API_BEGIN
    if (!butterflies_per_s) raise(HX_ERR_INVALID_ARGUMENT, "butterflies_per_s pointer must not be null.");
    *butterflies_per_s = butterfly_rate(as_stream(stream));
    HX_API_END
}

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

int hx_ntt_plan_workspace_bytes(const hx_ntt_plan* plan, const hx_view* a, const hx_view* result, size_t* bytes) {
    HX_API_BEGIN
    if (!bytes) raise(HX_ERR_INVALID_ARGUMENT, "bytes pointer must not be null.");
    *bytes = ntt_workspace_bytes(plan, a, result);
    HX_API_END
}

int hx_ntt_plan_exec_ws(const hx_ntt_plan* plan, const hx_view* a, const hx_view* result, void* workspace,
                        size_t workspace_bytes, hx_stream stream) {
    HX_API_BEGIN
    ntt_plan_exec_ws(plan, a, result, workspace, workspace_bytes, as_stream(stream));
    HX_API_END
}

int hx_ntt_plan_exec_noalloc(const hx_ntt_plan* plan, const hx_view* a, const hx_view* result, hx_stream stream) {
    HX_API_BEGIN
    ntt_plan_exec_noalloc(plan, a, result, as_stream(stream));
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
    // freed after the queued passes have run (stream-ordered)
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
