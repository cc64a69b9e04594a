"""Exact 65536-point negacyclic NTT on the INT8 matrix cores (primes < 2^28).

The device side (hx_ntt_mfma_* in include/heal_amd.h) runs the transform as
two 256 x 256 x 256 integer products per column; this module builds its
tables from the primes and their primitive 2N-th roots psi (the same psi that
define the standard twiddles psi^bitrev(j) of the radix-2 NTT), and wraps the
plan.  The result equals hx_ntt with those twiddles and the bit-reversal
permutation, bit for bit.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import capi

N = 65536
R = 256  # N = R * R


def _powers(base: int, q: int, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.int64)
    v = 1
    for i in range(count):
        out[i] = v
        v = v * base % q
    return out


def _planes(w: np.ndarray) -> np.ndarray:
    """[256 u][256 k] values < 2^28 -> uint8 [4 planes][16 u-tiles][4 k-chunks][64 lanes][16]:
    lane L holds row u = 16 ut + L mod 16 and k = 64 kq + 16 (L // 16) + i in byte i."""
    pl = np.stack([(w >> (7 * l)) & 127 for l in range(4)]).astype(np.uint8)  # [4][256][256]
    pl = pl.reshape(4, 16, 16, 4, 4, 16)                                    # l, ut, r, kq, g, i
    return np.ascontiguousarray(pl.transpose(0, 1, 3, 4, 2, 5)).reshape(4, 16, 4, 64, 16)


def tables(primes, psis):
    """Host tables of hx_ntt_mfma_plan_create for the given primes and psi."""
    k = len(primes)
    w2 = np.empty((k, 4, 16, 4, 64, 16), dtype=np.uint8)
    w1 = np.empty_like(w2)
    tm = np.empty((k, R, R), dtype=np.uint32)
    red = np.zeros((k, 10), dtype=np.uint64)
    u = np.arange(R, dtype=np.int64)
    for t, (q, psi) in enumerate(zip(primes, psis)):
        if q >= 1 << 28:
            raise ValueError("the matrix-core NTT needs primes below 2^28")
        if pow(psi, N, q) != q - 1:
            raise ValueError("psi must be a primitive 2N-th root of unity mod q")
        omega = psi * psi % q
        o256 = pow(omega, R, q)                       # primitive 256-th root
        p_o256 = _powers(o256, q, R)
        uu = np.outer(u, u)                           # u * i < 65536
        w1m = p_o256[uu % R]                          # omega^(256 u1 i1)
        psi256 = _powers(pow(psi, R, q), q, R)        # psi^(256 i2)
        w2m = w1m * psi256[None, :] % q               # omega^(256 u2 i2) psi^(256 i2)
        o_lo = _powers(omega, q, R)                   # omega^j, j < 256
        o_hi = _powers(pow(omega, R, q), q, R)        # omega^(256 j)
        om = o_hi[uu // R] * o_lo[uu % R] % q         # omega^(u2 i1)
        tmat = om * _powers(psi, q, R)[None, :] % q   # psi^i1 omega^(u2 i1)
        w2[t], w1[t], tm[t] = _planes(w2m), _planes(w1m), tmat.astype(np.uint32)
        red[t, 0] = q
        red[t, 1] = ((1 << 64) - 1) // q
        for s in range(7):
            red[t, 2 + s] = pow(2, 7 * s, q)
    return w2, w1, tm, red


class NttMfmaPlan:
    """The matrix-core NTT for primes q_t < 2^28 with roots psi_t (N = 65536)."""

    def __init__(self, primes, psis, stream=None):
        w2, w1, tm, red = tables(primes, psis)
        self.k = len(primes)
        self._h = ctypes.c_void_p()
        ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        capi.check(capi.lib.hx_ntt_mfma_plan_create(ctypes.byref(self._h), self.k, ptr(w2), ptr(w1), ptr(tm),
                                                    ptr(red), capi.stream_of() if stream is None else stream))

    def __call__(self, a, out):
        """a, out: int64 [l, 65536, 1, k] device views."""
        capi._need_cuda(a, out)
        capi.check(capi.lib.hx_ntt_mfma_exec(self._h, capi.view(a), capi.view(out), capi.stream_of(out)))
        return out

    def close(self):
        if self._h:
            capi.check(capi.lib.hx_ntt_mfma_plan_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def small_primes(count: int) -> list[int]:
    """The largest `count` primes q < 2^28 with q = 1 mod 2^17 (so 1 mod 2N)."""
    from .params import is_prime

    out, c = [], ((1 << 28) - 1) >> 17
    while len(out) < count and c > 0:
        q = (c << 17) + 1
        if q < (1 << 28) and is_prime(q):
            out.append(q)
        c -= 1
    return out


def _main():
    import argparse
    import json

    from .params import NttParams

    ap = argparse.ArgumentParser(description="matrix-core NTT vs the ALU NTT: 16 polynomials x 32 primes < 2^28")
    ap.add_argument("--polys", type=int, default=16)
    ap.add_argument("--limbs", type=int, default=32)
    ap.add_argument("--reps", type=int, default=10)
    args = ap.parse_args()
    primes = small_primes(args.limbs)
    prm = NttParams(N, primes, device="cuda")
    plan = NttMfmaPlan(primes, prm.psi)
    alu = capi.NttPlan(prm.p, prm.perm, prm.twiddles)
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    q = prm.p
    # limb-major storage [polys, limbs, N] viewed as [l, N, 1, k]
    data = torch.randint(0, 1 << 62, (args.polys, args.limbs, N), generator=g, device="cuda") % q.view(1, -1, 1)
    a = data.permute(0, 2, 1).unsqueeze(2)
    o1 = torch.empty_like(data).permute(0, 2, 1).unsqueeze(2)
    o2 = torch.empty_like(data).permute(0, 2, 1).unsqueeze(2)
    plan(a, o1)
    alu(a, o2)
    torch.cuda.synchronize()
    equal = bool(torch.equal(o1, o2))

    def timed(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        fn()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(args.reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1000.0 / args.reps

    t_mfma = timed(lambda: plan(a, o1))
    t_alu = timed(lambda: alu(a, o2))
    limbs = args.polys * args.limbs
    ops = limbs * 2 * 16 * 2 * R ** 3  # 2 steps x 16 plane products x 2 ops/MAC x 256^3 MACs
    print(json.dumps({"metric": f"NTT of {args.polys} x {args.limbs} limbs (N=65536, q < 2^28)",
                      "mfma_equal_alu": equal, "mfma_ms": round(t_mfma * 1e3, 4), "alu_ms": round(t_alu * 1e3, 4),
                      "mfma_limb_GBps": round(limbs * 2 * N * 8 / t_mfma / 1e9, 1),
                      "alu_limb_GBps": round(limbs * 2 * N * 8 / t_alu / 1e9, 1),
                      "mfma_int8_TOPS": round(ops / t_mfma / 1e12, 1), "int8_peak_TOPS": 5000.0,
                      "mfma_frac_of_int8_peak": round(ops / t_mfma / 5e15, 4)}), flush=True)


if __name__ == "__main__":
    _main()
