/*
 * heal_oracle.c — CPU restatement of the HEAL reference's hot-path operations.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product (heal_amd) never links, loads or calls it.
 *
 * Each function follows the loop structure of the reference file it cites and
 * uses the same double-width arithmetic (int32 -> int64, int64 -> __int128,
 * example_impl/typing.h:8-20), so on every input where the reference is
 * defined the output is the reference's output.  Views are hx_view descriptors
 * over HOST memory.  Status codes are those of include/heal_amd.h.
 *
 * Where the reference is undefined (signed overflow, division by zero, reading
 * outside a buffer) the oracle states the behaviour the product implements:
 *   - INTT scaling (example_impl/ntt_impl.cpp:165) multiplies in T and
 *     overflows for q > ~2^31.5 (int64); or_intt(mode=0) computes the exact
 *     scaling, mode=1 the two's-complement wrap that gcc emits (used only to
 *     compare with the compiled reference at those primes).
 *   - negative perm/perms entries read before the buffer in the reference; here
 *     they are HX_ERR_OUT_OF_RANGE like entries >= m.
 *   - a modulus or divisor of 0 gives 0.
 *   - base_bits >= 64 in g_decomposition is HX_ERR_INVALID_ARGUMENT.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/heal_amd.h"

typedef __int128 i128;

static char g_err[256];
const char* or_last_error(void) { return g_err; }

static int fail(int code, const char* msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return code;
}

static int64_t numel(const hx_view* v) {
    int64_t n = 1;
    for (int i = 0; i < v->ndim; ++i) n *= v->dims[i];
    return n;
}

static int64_t rd(const hx_view* v, int64_t off) {
    if (v->dtype == HX_INT32) return ((const int32_t*)v->data)[off];
    return ((const int64_t*)v->data)[off];
}

static void wr(const hx_view* v, int64_t off, int64_t x) {
    if (v->dtype == HX_INT32)
        ((int32_t*)v->data)[off] = (int32_t)x;
    else
        ((int64_t*)v->data)[off] = x;
}

/* cast to the element type T of view v (int32 truncation, as static_cast<T>) */
static int64_t to_t(const hx_view* v, int64_t x) { return v->dtype == HX_INT32 ? (int64_t)(int32_t)x : x; }

static int same_dims(const hx_view* a, const hx_view* b) {
    if (a->ndim != b->ndim) return 0;
    for (int i = 0; i < a->ndim; ++i)
        if (a->dims[i] != b->dims[i]) return 0;
    return 1;
}

/* linear index -> coordinates of a row-major shape */
static void unravel(int64_t idx, const int64_t* dims, int nd, int64_t* coord) {
    for (int d = nd - 1; d >= 0; --d) {
        coord[d] = dims[d] ? idx % dims[d] : 0;
        idx = dims[d] ? idx / dims[d] : 0;
    }
}

static int64_t offset_of(const hx_view* v, const int64_t* coord) {
    int64_t off = 0;
    for (int i = 0; i < v->ndim; ++i) off += coord[i] * v->strides[i];
    return off;
}

/* DeviceTensor::at_with_broadcast (example_impl/device_memory_impl.cpp:111-124):
 * right-aligned, size-1 dims index 0, other dims take the result coordinate and
 * must be in bounds (at() throws out_of_range otherwise). Returns -1 if OOB. */
static int64_t bcast_offset(const hx_view* v, const int64_t* full, int nfull) {
    int shift = nfull - v->ndim;
    int64_t off = 0;
    for (int i = 0; i < v->ndim; ++i) {
        int64_t c = v->dims[i] == 1 ? 0 : full[i + shift];
        if (c >= v->dims[i]) return -1;
        off += c * v->strides[i];
    }
    return off;
}

/* CHECK_DIMS_BROADCASTABLE (example_impl/modop_impl.cpp:89-101) */
static int check_broadcastable(const hx_view* t, const hx_view* r, const char* label) {
    static char msg[128];
    if (t->ndim > r->ndim) {
        snprintf(msg, sizeof msg, "%s has more dims than result.", label);
        return fail(HX_ERR_INVALID_ARGUMENT, msg);
    }
    for (int i = 1; i <= t->ndim; ++i) {
        int64_t td = t->dims[t->ndim - i], rdim = r->dims[r->ndim - i];
        if (td != 1 && td != rdim) {
            snprintf(msg, sizeof msg, "%s not broadcast-compatible with result.", label);
            return fail(HX_ERR_INVALID_ARGUMENT, msg);
        }
    }
    return HX_OK;
}

/* (T_DP(a) op T_DP(b)) % T_DP(p) with truncating %, example_impl/modop_impl.cpp:123-126, 238-239 */
static int64_t modop_value(int op, int is32, int64_t a, int64_t b, int64_t p) {
    if (p == 0) return 0;
    if (is32) {
        int64_t t = op == 0 ? a * b : a + b; /* int64 holds both exactly */
        return (int64_t)(int32_t)(t % p);
    }
    i128 t = op == 0 ? (i128)a * (i128)b : (i128)a + (i128)b;
    return (int64_t)(t % (i128)p);
}

/* modmul_* / modsum_* (example_impl/modop_impl.cpp:44-80, 108-180).
 * op: 0 = modmul, 1 = modsum.  b == NULL uses b_scalar; p == NULL uses p_scalar. */
int or_modop(int op, const hx_view* a, const hx_view* b, int64_t b_scalar, const hx_view* p, int64_t p_scalar,
             const hx_view* result) {
    int st;
    if ((st = check_broadcastable(a, result, "a"))) return st;
    if (b && (st = check_broadcastable(b, result, "b"))) return st;
    if (p) {
        /* CHECK_DIMS_MATCH_LAST (:84-87) */
        if (p->ndim != 1 && (p->ndim == 0 || result->ndim == 0 || p->dims[p->ndim - 1] != result->dims[result->ndim - 1]))
            return fail(HX_ERR_INVALID_ARGUMENT, "Last dimension of p must match last dimension of result.");
        if (p->ndim > result->ndim) return fail(HX_ERR_INVALID_ARGUMENT, "p has more dims than result.");
    }
    int is32 = result->dtype == HX_INT32;
    int nd = result->ndim;
    int64_t total = numel(result);
    int64_t coord[HX_MAX_DIMS];
    int64_t bs = to_t(result, b_scalar), ps = to_t(result, p_scalar);
    for (int64_t idx = 0; idx < total; ++idx) {
        unravel(idx, result->dims, nd, coord);
        int64_t av = rd(a, bcast_offset(a, coord, nd));
        int64_t bv = bs, pv = ps;
        if (b) bv = rd(b, bcast_offset(b, coord, nd));
        if (p) {
            int64_t po = bcast_offset(p, coord, nd);
            if (po < 0) return fail(HX_ERR_OUT_OF_RANGE, "Index out of bounds.");
            pv = rd(p, po);
        }
        wr(result, offset_of(result, coord), modop_value(op, is32, av, bv, pv));
    }
    return HX_OK;
}

/* mod_tt / mod_tc / mod_ct (example_impl/modop_impl.cpp:13-42, 183-235).
 * a == NULL uses a_scalar, b == NULL uses b_scalar; `nulls` flags which of a/b
 * were passed as null tensors (-> invalid_argument). */
int or_mod(const hx_view* a, int64_t a_scalar, const hx_view* b, int64_t b_scalar, const hx_view* result,
           int a_is_tensor, int b_is_tensor) {
    if (a_is_tensor && !a) return fail(HX_ERR_INVALID_ARGUMENT, "a pointer must not be null.");
    if (b_is_tensor && !b) return fail(HX_ERR_INVALID_ARGUMENT, "b pointer must not be null.");
    if (a && !same_dims(a, result)) return fail(HX_ERR_INVALID_ARGUMENT, "a must have exactly the same shape as result.");
    if (b && !same_dims(b, result)) return fail(HX_ERR_INVALID_ARGUMENT, "b must have exactly the same shape as result.");
    int nd = result->ndim;
    int64_t total = numel(result), coord[HX_MAX_DIMS];
    int64_t as = to_t(result, a_scalar), bs = to_t(result, b_scalar);
    for (int64_t idx = 0; idx < total; ++idx) {
        unravel(idx, result->dims, nd, coord);
        int64_t av = a ? rd(a, offset_of(a, coord)) : as;
        int64_t bv = b ? rd(b, offset_of(b, coord)) : bs;
        int64_t r;
        if (bv == 0 || (bv == -1))
            r = 0; /* x % 0 is undefined; x % -1 is 0 (INT_MIN % -1 undefined) */
        else
            r = av % bv;
        wr(result, offset_of(result, coord), r);
    }
    return HX_OK;
}

/* axis_modsum (example_impl/axis_modsum_impl.cpp:12-76): sequential
 * sum = (sum + a) % mod in T, mod = p[last coordinate]. */
int or_axis_modsum(const hx_view* a, const hx_view* p, const hx_view* result, int64_t axis) {
    if (p->ndim != 1) return fail(HX_ERR_INVALID_ARGUMENT, "p must be a 1D tensor of shape [k]");
    int nd = a->ndim;
    if (axis < 0 || axis >= nd - 1)
        return fail(HX_ERR_INVALID_ARGUMENT, "axis must be in range [0, ndim - 2] (can't reduce across last axis)");
    if (a->dims[nd - 1] != p->dims[0]) return fail(HX_ERR_INVALID_ARGUMENT, "Last dimension of a must match shape of p");
    if (result->ndim != nd - 1) return fail(HX_ERR_INVALID_ARGUMENT, "result must have the shape of a without axis");
    for (int i = 0, j = 0; i < nd; ++i) {
        if (i == axis) continue;
        if (result->dims[j++] != a->dims[i]) return fail(HX_ERR_INVALID_ARGUMENT, "result must have the shape of a without axis");
    }
    int64_t total = numel(result), rc[HX_MAX_DIMS], ic[HX_MAX_DIMS];
    for (int64_t idx = 0; idx < total; ++idx) {
        unravel(idx, result->dims, nd - 1, rc);
        for (int i = 0, j = 0; i < nd; ++i) ic[i] = (i == axis) ? 0 : rc[j++];
        int64_t mod = rd(p, ic[nd - 1] * p->strides[0]);
        int64_t sum = 0;
        for (int64_t r = 0; r < a->dims[axis]; ++r) {
            ic[axis] = r;
            int64_t s = sum + rd(a, offset_of(a, ic));
            sum = mod == 0 ? 0 : to_t(a, s % mod);
        }
        wr(result, offset_of(result, rc), sum);
    }
    return HX_OK;
}

/* g_decomposition (example_impl/g_decomposition_impl.cpp:10-67): unsigned
 * digits (`base` is size_t), value /= base assigned back to T. */
int or_g_decomposition(const hx_view* a, const hx_view* result, uint64_t power, uint64_t base_bits,
                       uint64_t* overflow) {
    int ok = result->ndim == a->ndim + 1 && result->dims[result->ndim - 1] == (int64_t)power;
    for (int i = 0; ok && i < a->ndim; ++i) ok = result->dims[i] == a->dims[i];
    if (!ok) return fail(HX_ERR_INVALID_ARGUMENT, "Output must have shape a.shape + [power]");
    if (base_bits >= 64) return fail(HX_ERR_INVALID_ARGUMENT, "base_bits must be < 64");
    uint64_t base = 1ull << base_bits;
    int64_t total = numel(a), c[HX_MAX_DIMS + 1];
    uint64_t over = 0;
    for (int64_t idx = 0; idx < total; ++idx) {
        unravel(idx, a->dims, a->ndim, c);
        int64_t value = rd(a, offset_of(a, c));
        for (uint64_t d = 0; d < power; ++d) {
            c[a->ndim] = (int64_t)d;
            uint64_t u = (uint64_t)value;
            wr(result, offset_of(result, c), to_t(a, (int64_t)(u % base)));
            value = to_t(a, (int64_t)(u / base));
        }
        if (value > 0) ++over;
    }
    if (overflow) *overflow = over;
    return HX_OK;
}

/* validate_ntt_inputs (example_impl/ntt_impl.cpp:14-43) */
static int ntt_validate(const hx_view* a, const hx_view* p, const hx_view* perm, const hx_view* tw,
                        const hx_view* result) {
    if (a->ndim != 4) return fail(HX_ERR_INVALID_ARGUMENT, "Input tensor 'a' must have shape [l, m, r, k].");
    if (!same_dims(result, a)) return fail(HX_ERR_INVALID_ARGUMENT, "Output tensor must have the same shape as input tensor.");
    if (p->ndim != 1 || p->dims[0] != a->dims[3]) return fail(HX_ERR_INVALID_ARGUMENT, "Tensor 'p' must have shape [k].");
    if (perm->ndim != 1 || perm->dims[0] != a->dims[1]) return fail(HX_ERR_INVALID_ARGUMENT, "Tensor 'perm' must have shape [m].");
    if (tw->ndim != 2 || tw->dims[0] != a->dims[3] || tw->dims[1] != a->dims[1])
        return fail(HX_ERR_INVALID_ARGUMENT, "Tensor 'twiddles' must have shape [k, m].");
    return HX_OK;
}

static int64_t at4(const hx_view* v, int64_t i, int64_t u, int64_t j, int64_t t) {
    return i * v->strides[0] + u * v->strides[1] + j * v->strides[2] + t * v->strides[3];
}

/* T_DP product reduced by truncating %, cast to T */
static int64_t mulmod_t(int is32, int64_t x, int64_t y, int64_t q) {
    if (is32) return (int64_t)(int32_t)((x * y) % q);
    return (int64_t)(((i128)x * (i128)y) % (i128)q);
}

/* ntt (example_impl/ntt_impl.cpp:69-116): per (i, j, t), copy the column,
 * radix-2 Cooley-Tukey with s = tw[t, stage + u], then out[u] = x[perm[u]]
 * (apply_permutation, :45-65). */
int or_ntt(const hx_view* a, const hx_view* p, const hx_view* perm, const hx_view* tw, const hx_view* result) {
    int st = ntt_validate(a, p, perm, tw, result);
    if (st) return st;
    int64_t l = a->dims[0], m = a->dims[1], r = a->dims[2], k = a->dims[3];
    int is32 = a->dtype == HX_INT32;
    int64_t* x = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m ? m : 1));
    int64_t* tmp = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m ? m : 1));
    for (int64_t u = 0; u < m; ++u) {
        int64_t pu = rd(perm, u * perm->strides[0]);
        if (pu < 0 || pu >= m) {
            free(x);
            free(tmp);
            return fail(HX_ERR_OUT_OF_RANGE, "Index out of bounds.");
        }
    }
    for (int64_t i = 0; i < l; ++i)
        for (int64_t j = 0; j < r; ++j)
            for (int64_t t = 0; t < k; ++t) {
                int64_t mod = rd(p, t * p->strides[0]);
                for (int64_t u = 0; u < m; ++u) x[u] = rd(a, at4(a, i, u, j, t));
                int64_t step = m;
                for (int64_t stage = 1; stage < m; stage *= 2) {
                    step /= 2;
                    for (int64_t u = 0; u < stage; ++u) {
                        int64_t j1 = 2 * u * step, j2 = j1 + step;
                        int64_t s = rd(tw, t * tw->strides[0] + (stage + u) * tw->strides[1]);
                        for (int64_t jx = j1; jx < j2; ++jx) {
                            int64_t uv = x[jx];
                            int64_t vm = mulmod_t(is32, x[jx + step], s, mod);
                            x[jx] = to_t(a, (uv + vm) % mod);
                            x[jx + step] = to_t(a, (uv + mod - vm) % mod);
                        }
                    }
                }
                for (int64_t u = 0; u < m; ++u) tmp[u] = x[rd(perm, u * perm->strides[0])];
                for (int64_t u = 0; u < m; ++u) wr(result, at4(result, i, u, j, t), tmp[u]);
            }
    free(x);
    free(tmp);
    return HX_OK;
}

/* intt (example_impl/ntt_impl.cpp:118-170): scatter x[perm[u]] = a[u] into the
 * result column (positions not hit keep the result's previous contents),
 * Gentleman-Sande with s = itw[t, half + group], then multiply by m_inv[t].
 * mode 0: exact scaling; mode 1: the reference's T-width product (wraps). */
int or_intt(const hx_view* a, const hx_view* p, const hx_view* perm, const hx_view* itw, const hx_view* minv,
            const hx_view* result, int mode) {
    int st = ntt_validate(a, p, perm, itw, result);
    if (st) return st;
    int64_t l = a->dims[0], m = a->dims[1], r = a->dims[2], k = a->dims[3];
    int is32 = a->dtype == HX_INT32;
    if (minv->ndim != 1) return fail(HX_ERR_INVALID_ARGUMENT, "Number of indices does not match tensor dimensions.");
    if (minv->dims[0] < k) return fail(HX_ERR_OUT_OF_RANGE, "Index out of bounds.");
    for (int64_t u = 0; u < m; ++u) {
        int64_t pu = rd(perm, u * perm->strides[0]);
        if (pu < 0 || pu >= m) return fail(HX_ERR_OUT_OF_RANGE, "Index out of bounds.");
    }
    int64_t* x = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m ? m : 1));
    int64_t* av = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m ? m : 1));
    for (int64_t i = 0; i < l; ++i)
        for (int64_t j = 0; j < r; ++j)
            for (int64_t t = 0; t < k; ++t) {
                int64_t mod = rd(p, t * p->strides[0]);
                int64_t mi = rd(minv, t * minv->strides[0]);
                /* read the whole input column first so that a == result behaves
                 * like distinct buffers */
                for (int64_t u = 0; u < m; ++u) {
                    av[u] = rd(a, at4(a, i, u, j, t));
                    x[u] = rd(result, at4(result, i, u, j, t));
                }
                for (int64_t u = 0; u < m; ++u) x[rd(perm, u * perm->strides[0])] = av[u];
                int64_t half = m / 2, ts = 1;
                while (half >= 1) {
                    for (int64_t tid = 0; tid < m / 2; ++tid) {
                        int64_t group = tid / ts;
                        int64_t iu = group * ts * 2 + (tid % ts), iv = iu + ts;
                        int64_t s = rd(itw, t * itw->strides[0] + (half + group) * itw->strides[1]);
                        int64_t uv = x[iu], vv = x[iv];
                        x[iu] = to_t(a, (uv + vv) % mod);
                        x[iv] = mulmod_t(is32, to_t(a, uv + mod - vv), s, mod);
                    }
                    ts *= 2;
                    half /= 2;
                }
                for (int64_t u = 0; u < m; ++u) {
                    int64_t val = x[u];
                    int64_t out;
                    if (mode == 1) {
                        int64_t prod = is32 ? (int64_t)(int32_t)((uint32_t)val * (uint32_t)mi)
                                            : (int64_t)((uint64_t)val * (uint64_t)mi);
                        out = to_t(a, prod % mod);
                    } else {
                        out = mulmod_t(is32, val, mi, mod);
                    }
                    wr(result, at4(result, i, u, j, t), out);
                }
            }
    free(x);
    free(av);
    return HX_OK;
}

/* permute (example_impl/permute_impl.cpp:8-65) */
int or_permute(const hx_view* a, const hx_view* perms, const hx_view* result, int64_t ew, int64_t pa) {
    int nd = a->ndim;
    if (!same_dims(result, a)) return fail(HX_ERR_INVALID_ARGUMENT, "Result shape must match input shape.");
    if (ew < 0 || ew >= nd || pa < 0 || pa >= nd) return fail(HX_ERR_INVALID_ARGUMENT, "Axis out of bounds.");
    if (ew == pa) return fail(HX_ERR_INVALID_ARGUMENT, "elementwise_axis and perm_axis must be different.");
    int64_t l = a->dims[ew], m = a->dims[pa];
    if (perms->ndim != 2 || perms->dims[0] != l || perms->dims[1] != m)
        return fail(HX_ERR_INVALID_ARGUMENT,
                    "Perms must have shape [l, m] where l and m match a.shape at elementwise and perm axes.");
    int64_t total = numel(a), c[HX_MAX_DIMS];
    /* gather into a scratch copy so that result == a behaves like distinct buffers */
    int64_t* out = (int64_t*)malloc(sizeof(int64_t) * (size_t)(total ? total : 1));
    for (int64_t idx = 0; idx < total; ++idx) {
        unravel(idx, a->dims, nd, c);
        int64_t pi = rd(perms, c[ew] * perms->strides[0] + c[pa] * perms->strides[1]);
        if (pi < 0 || pi >= m) {
            free(out);
            return fail(HX_ERR_OUT_OF_RANGE, "Permutation index out of bounds.");
        }
        int64_t keep = c[pa];
        c[pa] = pi;
        out[idx] = rd(a, offset_of(a, c));
        c[pa] = keep;
    }
    for (int64_t idx = 0; idx < total; ++idx) {
        unravel(idx, a->dims, nd, c);
        wr(result, offset_of(result, c), out[idx]);
    }
    free(out);
    return HX_OK;
}

/* make_contiguous gather (example_impl/contiguous_impl.cpp:31-40), any element size */
int or_strided_copy(const hx_view* src, const hx_view* dst) {
    if (!same_dims(src, dst)) return fail(HX_ERR_INVALID_ARGUMENT, "src and dst must have the same shape");
    size_t es = src->dtype == HX_INT32 ? 4 : 8;
    int64_t total = numel(src), c[HX_MAX_DIMS];
    for (int64_t idx = 0; idx < total; ++idx) {
        unravel(idx, src->dims, src->ndim, c);
        memcpy((char*)dst->data + (size_t)offset_of(dst, c) * es, (const char*)src->data + (size_t)offset_of(src, c) * es, es);
    }
    return HX_OK;
}
