// hx_ntt_mfma.hip — exact negacyclic NTT of length 65536 for primes q < 2^28 as
// a four-step transform whose 256-point sub-transforms are INT8 matrix-core
// products (v_mfma_i32_16x16x64_i8 on gfx950).
//
// The transform is the reference's NTT with the standard tables
// (example_impl/ntt_impl.cpp:69-116 with twiddles[t, j] = psi^bitrev(j) and
// perm = bitrev): out[u] = sum_i a[i] psi^((2u+1) i) mod q.  With
// omega = psi^2, i = i1 + 256 i2 and u = u2 + 256 u1:
//   step 1  C[u2, i1]  = sum_i2 W2[u2, i2] a[256 i2 + i1],  W2 = omega^(256 u2 i2) psi^(256 i2)
//           C'[u2, i1] = C[u2, i1] * T[u2, i1] mod q,       T  = psi^i1 omega^(u2 i1)
//   step 2  out[256 u1 + u2] = sum_i1 W1[u1, i1] C'[u2, i1], W1 = omega^(256 u1 i1)
// Both steps are 256 x 256 x 256 products per column with a constant left
// matrix.  Operands (< 2^28) are split into four 7-bit int8 planes; the 16
// plane products accumulate in int32 grouped by s = l + k (each group sums at
// most 4 * 256 * 127^2 < 2^24), and a value is recombined as
// sum_s acc_s * (2^(7 s) mod q) < 2^55, then reduced by Barrett with
// mu = floor((2^64 - 1) / q): for v < 2^56 the estimate floor(v mu / 2^64) is
// floor(v / q) or one less, so one conditional subtraction finishes.  Every
// step is exact, so the output equals the radix-2 network's bit for bit.
//
// MFMA operand layout (checked on the device): lane l supplies row l mod 16 of
// A and column l mod 16 of B with 16 int8 each, k = 16 (l / 16) + i within a
// 64-wide k chunk (A and B agree); the 16 x 16 int32 result gives lane l rows
// 4 (l / 16) + j, j < 4, of column l mod 16.
//
// Step 1 computes the transpose C^T = a^T W2^T (the data planes as A, the W2
// planes as B: the two operand layouts agree, so the same registers serve).
// A lane then holds C'[u2, i1] for four consecutive i1 of one u2, which are the
// four bytes of one word of step 2's right operand (k = i1, column u2), and
// step 1 stores C' as step 2's int8 planes in step 2's LDS operand order
// (`mid` below); step 2 copies its operand block with 16-byte loads.
#include <cstring>
#include <new>

#include "hx_internal.hpp"

struct hx_ntt_mfma_plan {
    int device;
    int64_t k;      // limbs
    void* mem;      // tables (below)
    uint4* w2;      // [k][4 planes][16 u-tiles][4 k-chunks][64 lanes]
    uint4* w1;      // same for W1
    uint32_t* tmat; // [k][256 u2][256 i1]
    uint64_t* red;  // [k][10]: q, mu, R_0..R_6, pad
    hx::Workspaces* wss;  // the step-1 intermediate, per stream (see hx::Workspaces)
};

namespace hx {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kRows = 256;
constexpr int kBlockCols = 32;            // columns of the right operand per workgroup
constexpr int kColBlocks = kRows / kBlockCols;
constexpr int kRowHalves = 2;             // a workgroup computes 128 of the 256 product rows
constexpr int kThreads4 = 512;            // 8 waves: 16 rows x 32 columns of the product each

struct Ntt4Args {
    const int64_t* a;      // input element (p, i, t) at a + p*a_sp + t*a_st + i*a_si
    int64_t a_sp, a_st, a_si;
    int64_t* out;
    int64_t o_sp, o_st, o_si;
    uint32_t* mid;         // step 2's right operand planes: per column [8 c-blocks][2 c-tiles][4 planes][4 k-chunks][64 lanes] x 16 bytes
    const uint4* w;        // the step's left matrix planes
    const uint32_t* tmat;
    const uint64_t* red;
    int64_t P, K;
};

__device__ __forceinline__ uint64_t barrett(uint64_t v, uint64_t q, uint64_t mu) {
    const uint64_t qt = __umul64hi(v, mu);
    uint64_t r = v - qt * q;
    return r >= q ? r - q : r;
}

// residue in [0, q) of a signed input value (inputs outside [0, q) are rare:
// a plain 64-bit remainder)
__device__ __forceinline__ uint64_t reduce_in(int64_t x, uint64_t q) {
    if (x >= 0 && (uint64_t)x < q) return (uint64_t)x;
    const uint64_t ax = x < 0 ? (uint64_t)0 - (uint64_t)x : (uint64_t)x;
    const uint64_t r = ax % q;
    return x < 0 ? (r ? q - r : 0) : r;
}

template <int STEP>
__global__ __launch_bounds__(kThreads4) void k_ntt4(Ntt4Args A) {
    __shared__ uint4 bs[2 * 4 * 4 * 64];  // right operand planes: [c-tile][plane][k-chunk][lane]
    // workgroups of one limb are adjacent (they share the left matrix in L2), and
    // so are the two row halves of one column block (they share its right operand)
    // (blocks are dispatched to the 8 XCDs round-robin: give each XCD a contiguous range)
    const int64_t nblk = A.P * A.K * kColBlocks * kRowHalves;
    const int64_t xcd = (int64_t)blockIdx.x % 8, rem = nblk % 8;
    const int64_t blk = xcd * (nblk / 8) + (xcd < rem ? xcd : rem) + (int64_t)blockIdx.x / 8;
    const int64_t t = blk / (A.P * kColBlocks * kRowHalves);
    const int64_t p = (blk / (kColBlocks * kRowHalves)) % A.P;
    const int cb = (int)((blk / kRowHalves) % kColBlocks);
    const int rh = (int)(blk % kRowHalves);
    const int64_t col = p * A.K + t;
    const uint64_t* rd = A.red + t * 10;
    const uint64_t q = rd[0], mu = rd[1];
    const int tid = threadIdx.x;

    // ---- the data operand: 256 k rows x 32 columns as int8 planes in LDS
    if (STEP == 1) {
        // split here: a[256 i2 + i1], k = i2, column i1
        const int c = tid % kBlockCols, run = tid / kBlockCols;  // 16 runs of 16 k
        const int gc = cb * kBlockCols + c;
        uint32_t v[16];
        const int64_t* base = A.a + p * A.a_sp + t * A.a_st;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int64_t k = 16 * run + i;  // i2
            v[i] = (uint32_t)reduce_in(base[(k * kRows + gc) * A.a_si], q);
        }
        const int ct = c / 16, lane = (run % 4) * 16 + c % 16, kq = run / 4;
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            uint32_t w4[4];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                uint32_t word = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j) word |= ((v[4 * b + j] >> (7 * l)) & 127u) << (8 * j);
                w4[b] = word;
            }
            bs[((ct * 4 + l) * 4 + kq) * 64 + lane] = make_uint4(w4[0], w4[1], w4[2], w4[3]);
        }
    } else {
        // step 1 stored this column block's planes in this LDS order
        const uint4* src = (const uint4*)A.mid + (col * kColBlocks + cb) * (2 * 4 * 4 * 64);
#pragma unroll
        for (int i = 0; i < 4; ++i) bs[tid + i * kThreads4] = src[tid + i * kThreads4];
    }
    __syncthreads();

    // ---- 16 rows x 32 columns per wave: 1 x 2 tiles, 7 accumulators each
    const int wave = tid / 64, lane = tid % 64;
    const int ut = rh * 8 + wave;  // the wave's 16-row tile of the product
    v4i acc[2][7];
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int s = 0; s < 7; ++s) acc[b][s] = v4i{0, 0, 0, 0};
    const uint4* wl = A.w + t * (4 * 16 * 4 * 64);
    v4i am[4];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        const uint4 x = wl[((l * 16 + ut) * 4 + 0) * 64 + lane];
        am[l] = v4i{(int)x.x, (int)x.y, (int)x.z, (int)x.w};
    }
#pragma unroll
    for (int kq = 0; kq < 4; ++kq) {
        __builtin_amdgcn_sched_barrier(0);  // one k chunk of operands live at a time
        // the next chunk's left operand is requested before this chunk's products
        v4i an[4];
        if (kq + 1 < 4) {
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                const uint4 x = wl[((l * 16 + ut) * 4 + kq + 1) * 64 + lane];
                an[l] = v4i{(int)x.x, (int)x.y, (int)x.z, (int)x.w};
            }
        }
        v4i bm[2][4];
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                const uint4 x = bs[((ct * 4 + l) * 4 + kq) * 64 + lane];
                bm[ct][l] = v4i{(int)x.x, (int)x.y, (int)x.z, (int)x.w};
            }
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int la = 0; la < 4; ++la)
#pragma unroll
                for (int lb = 0; lb < 4; ++lb)
                    acc[ct][la + lb] = STEP == 1 ? __builtin_amdgcn_mfma_i32_16x16x64_i8(bm[ct][lb], am[la], acc[ct][la + lb], 0, 0, 0)
                                                 : __builtin_amdgcn_mfma_i32_16x16x64_i8(am[la], bm[ct][lb], acc[ct][la + lb], 0, 0, 0);
        if (kq + 1 < 4) {
#pragma unroll
            for (int l = 0; l < 4; ++l) am[l] = an[l];
        }
    }

    __builtin_amdgcn_sched_barrier(0);
    // ---- recombine, reduce, and write
    uint64_t R[7];
#pragma unroll
    for (int s = 0; s < 7; ++s) R[s] = rd[2 + s];
    if (STEP == 1) {
        // lane: C'[u2, i1] for u2 = ut*16 + lane%16 and i1 = i0 + j, j < 4, of each c-tile
        const int u2 = ut * 16 + lane % 16;
        const uint32_t* trow = A.tmat + (t * kRows + u2) * kRows;  // T[u2][*]
        uint32_t* planes = A.mid + (int64_t)col * (kRows * kRows);
        // step 2's operand entry of (u2, i1): c-block u2/32, c-tile (u2/16)%2, k-chunk i1/64,
        // lane ((i1%64)/16)*16 + u2%16, word (i1%16)/4, byte i1%4
        const int cb2 = u2 >> 5, ct2 = (u2 >> 4) & 1, kq2 = cb >> 1;
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            const int i0 = cb * kBlockCols + ct * 16 + 4 * (lane / 16);
            const uint4 tw4 = *(const uint4*)(trow + i0);
            const uint32_t tws[4] = {tw4.x, tw4.y, tw4.z, tw4.w};
            uint32_t r4[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                uint64_t v = 0;
#pragma unroll
                for (int s = 0; s < 7; ++s) v += (uint64_t)(uint32_t)acc[ct][s][j] * R[s];
                v = barrett(v, q, mu);
                r4[j] = (uint32_t)barrett(v * tws[j], q, mu);
            }
            const int lane2 = (((cb & 1) * 2 + ct) << 4) + lane % 16;
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                uint32_t word = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j) word |= ((r4[j] >> (7 * l)) & 127u) << (8 * j);
                const int64_t e = ((((int64_t)cb2 * 2 + ct2) * 4 + l) * 4 + kq2) * 64 + lane2;  // 16-byte entry
                planes[e * 4 + lane / 16] = word;
            }
        }
    } else {
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            const int row0 = ut * 16 + 4 * (lane / 16);
            const int gc = cb * kBlockCols + ct * 16 + lane % 16;
            uint32_t r4[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                uint64_t v = 0;
#pragma unroll
                for (int s = 0; s < 7; ++s) v += (uint64_t)(uint32_t)acc[ct][s][j] * R[s];
                r4[j] = (uint32_t)barrett(v, q, mu);
            }
            // out[256 u1 + u2], u1 = row, u2 = gc
            int64_t* ob = A.out + p * A.o_sp + t * A.o_st;
#pragma unroll
            for (int j = 0; j < 4; ++j) ob[((int64_t)(row0 + j) * kRows + gc) * A.o_si] = (int64_t)r4[j];
        }
    }
}

}  // namespace

static void require_ntt4_view(const hx_view* v, const char* name, int64_t k) {
    check_view(v, name);
    if (v->dtype != HX_INT64) raise(HX_ERR_INVALID_ARGUMENT, std::string(name) + " must be int64");
    if (v->ndim != 4 || v->dims[1] != 65536 || v->dims[2] != 1 || v->dims[3] != k)
        raise(HX_ERR_INVALID_ARGUMENT, std::string(name) + " must have shape [l, 65536, 1, k]");
}

hx_ntt_mfma_plan* ntt_mfma_plan_create(int64_t k, const void* w2, const void* w1, const void* tmat, const void* red,
                                       hipStream_t st) {
    if (k < 1) raise(HX_ERR_INVALID_ARGUMENT, "k must be positive");
    if (!w2 || !w1 || !tmat || !red) raise(HX_ERR_INVALID_ARGUMENT, "table pointers must not be null.");
    const size_t wb = (size_t)k * 4 * 16 * 4 * 64 * 16, tb = (size_t)k * kRows * kRows * 4, rb = (size_t)k * 10 * 8;
    auto* plan = new hx_ntt_mfma_plan;
    std::memset(plan, 0, sizeof *plan);
    plan->k = k;
    check_hip(hipGetDevice(&plan->device), "hipGetDevice");
    hipError_t e = hipMallocAsync(&plan->mem, 2 * wb + tb + rb, st);
    if (e != hipSuccess) {
        delete plan;
        check_hip(e, "hipMallocAsync(mfma ntt plan)");
    }
    char* m = (char*)plan->mem;
    plan->w2 = (uint4*)m;
    plan->w1 = (uint4*)(m + wb);
    plan->tmat = (uint32_t*)(m + 2 * wb);
    plan->red = (uint64_t*)(m + 2 * wb + tb);
    try {
        check_hip(hipMemcpyAsync(plan->w2, w2, wb, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
        check_hip(hipMemcpyAsync(plan->w1, w1, wb, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
        check_hip(hipMemcpyAsync(plan->tmat, tmat, tb, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
        check_hip(hipMemcpyAsync(plan->red, red, rb, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
        check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
        plan->wss = new Workspaces;
        plan->wss->device = plan->device;
    } catch (...) {
        (void)hipFreeAsync(plan->mem, st);
        (void)hipStreamSynchronize(st);
        delete plan;
        throw;
    }
    return plan;
}

void ntt_mfma_exec(const hx_ntt_mfma_plan* plan, const hx_view* a, const hx_view* out, hipStream_t st) {
    if (!plan) raise(HX_ERR_INVALID_ARGUMENT, "plan pointer must not be null.");
    require_ntt4_view(a, "a", plan->k);
    require_ntt4_view(out, "result", plan->k);
    if (a->dims[0] != out->dims[0]) raise(HX_ERR_INVALID_ARGUMENT, "Output tensor must have the same shape as input tensor.");
    int dev = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (dev != plan->device) raise(HX_ERR_INVALID_ARGUMENT, "plan belongs to another device");
    const int64_t P = a->dims[0], K = plan->k;
    if (P == 0) return;
    const size_t bytes = (size_t)P * K * kRows * kRows * 4;
    if (views_overlap(a, out) && !(a->data == out->data)) raise(HX_ERR_INVALID_ARGUMENT, "a and result overlap");
    std::lock_guard<std::mutex> lk(plan->wss->mu);
    const Workspaces::Ws w = plan->wss->acquire(st, 0, bytes);
    Ntt4Args A;
    std::memset(&A, 0, sizeof A);
    A.a = (const int64_t*)a->data;
    A.a_sp = a->strides[0], A.a_st = a->strides[3], A.a_si = a->strides[1];
    A.out = (int64_t*)out->data;
    A.o_sp = out->strides[0], A.o_st = out->strides[3], A.o_si = out->strides[1];
    A.mid = (uint32_t*)w.data;
    A.tmat = plan->tmat;
    A.red = plan->red;
    A.P = P;
    A.K = K;
    const int64_t blocks = P * K * kColBlocks * kRowHalves;
    if (blocks > 0x7fffffff) raise(HX_ERR_INVALID_ARGUMENT, "too many columns for one launch");
    A.w = plan->w2;
    hipLaunchKernelGGL(k_ntt4<1>, dim3((unsigned)blocks), dim3(kThreads4), 0, st, A);
    check_hip(hipGetLastError(), "mfma ntt step 1 launch");
    A.w = plan->w1;
    hipLaunchKernelGGL(k_ntt4<2>, dim3((unsigned)blocks), dim3(kThreads4), 0, st, A);
    check_hip(hipGetLastError(), "mfma ntt step 2 launch");
    plan->wss->note(st);
}

// freed after the work already queued with the plan (stream-ordered, no host wait)
void ntt_mfma_plan_destroy(hx_ntt_mfma_plan* plan) {
    if (!plan) return;
    std::vector<hipEvent_t> events;
    std::vector<void*> ptrs{plan->mem};
    {
        std::lock_guard<std::mutex> lk(plan->wss->mu);
        plan->wss->release(events, ptrs);
    }
    release_after(plan->device, events, ptrs);
    delete plan->wss;
    delete plan;
}

}  // namespace hx

using namespace hx;

extern "C" {

int hx_ntt_mfma_plan_create(hx_ntt_mfma_plan** out, int64_t k, const void* w2, const void* w1, const void* tmat,
                            const void* red, hx_stream stream) {
    HX_API_BEGIN
    if (!out) raise(HX_ERR_INVALID_ARGUMENT, "out pointer must not be null.");
    *out = nullptr;
    *out = ntt_mfma_plan_create(k, w2, w1, tmat, red, as_stream(stream));
    HX_API_END
}

int hx_ntt_mfma_exec(const hx_ntt_mfma_plan* plan, const hx_view* a, const hx_view* result, hx_stream stream) {
    HX_API_BEGIN
    ntt_mfma_exec(plan, a, result, as_stream(stream));
    HX_API_END
}

int hx_ntt_mfma_plan_destroy(hx_ntt_mfma_plan* plan) {
    HX_API_BEGIN
    ntt_mfma_plan_destroy(plan);
    HX_API_END
}

}  // extern "C"
