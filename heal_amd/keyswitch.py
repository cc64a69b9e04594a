"""Hybrid key-switch composed from HEAL calls on the MI355X backend.

The reference has no key-switch function; its API composes one as
  g_decomposition -> ntt -> modmul_ttt -> axis_modsum -> intt -> permute
with the limb axis last (x [N, L] -> digits [N, L, dnum] -> digit NTTs
[1, N, dnum, L] -> products with the evaluation key [dnum, 2, N, L] -> sum
over digits [2, N, L] -> coefficients -> Galois permutation).

`KeySwitchBatch` holds B independent instances (own x and own evaluation key
each) and runs the chain as six launches over the whole batch, using the
batch as the NTT's `l` axis; with B = 1 it is the plain per-call chain.
Plans and the validated Galois map are reused across calls, which is what a
caller that replays the same constants does.  `capture()` records the chain
into a hipGraph for replay.
"""
from __future__ import annotations

import torch

from . import capi
from . import params as P


class KeySwitchBatch:
    def __init__(self, n: int = 65536, limbs: int = 24, dnum: int = 3, base_bits: int = 60, batch: int = 1,
                 galois_elt: int = 5, seed: int = 0, device: str = "cuda", primes=None):
        self.n, self.L, self.dnum, self.w, self.B = n, limbs, dnum, base_bits, batch
        self.device = torch.device(device)
        self.primes = list(primes) if primes is not None else P.keyswitch_primes(limbs, n)
        self.prm = P.NttParams(n, self.primes, device=self.device)
        g = torch.Generator().manual_seed(seed)
        B, N, L, D = batch, n, limbs, dnum
        self.x = P.uniform_residues((B,), self.primes, N, g, self.device)                 # [B, N, L]
        self.evk = P.uniform_residues((D, B, 2), self.primes, N, g, self.device)          # [D, B, 2, N, L]
        self.perms = P.galois_perms(N, galois_elt, L, domain="coeff", device=self.device)  # [L, N]
        dev = self.device
        self.digits = torch.empty((B, N, L, D), dtype=torch.int64, device=dev)
        self.dntt = torch.empty((B, N, D, L), dtype=torch.int64, device=dev)
        self.prod = torch.empty((D, B, 2, N, L), dtype=torch.int64, device=dev)
        self.acc = torch.empty((B, 2, N, L), dtype=torch.int64, device=dev)
        self.coef = torch.empty((B * 2, N, 1, L), dtype=torch.int64, device=dev)
        self.out = torch.empty((B, 2, N, L), dtype=torch.int64, device=dev)
        self.fwd = capi.NttPlan(self.prm.p, self.prm.perm, self.prm.twiddles)
        self.inv = capi.NttPlan(self.prm.p, self.prm.perm, self.prm.inv_twiddles, self.prm.m_inv, inverse=True)
        # validate the Galois map once (synchronous, raises IndexError if out of range)
        capi.permute(self.coef.view(B * 2, N, L), self.perms, self.out.view(B * 2, N, L), 2, 1)
        self.graph = None
        self.fused = None
        self.key = None

    # -- fused form: three kernels, key in device layout ----------------------
    def prepare_fused(self):
        if self.fused is None:
            self.fused = capi.KeySwitchPlan(self.fwd, self.inv, self.perms, self.dnum, self.w)
            self.key = self.fused.prepare_key(self.evk)
        return self.fused

    def run_fused(self, out=None):
        self.prepare_fused()
        out = self.out if out is None else out
        self.fused(self.x, self.key, out)
        return out

    # -- the six HEAL calls -------------------------------------------------
    def run(self):
        B, N, L, D = self.B, self.n, self.L, self.dnum
        p = self.prm.p
        capi.g_decomposition(self.x, self.digits, D, self.w)
        self.fwd(self.digits.permute(0, 1, 3, 2), self.dntt)                       # [B, N, D, L] view
        dv = self.dntt.permute(2, 0, 1, 3).unsqueeze(2)                            # [D, B, 1, N, L] view
        capi.modmul_ttt(dv, self.evk, p, self.prod)
        capi.axis_modsum(self.prod, p, self.acc, 0)
        self.inv(self.acc.view(B * 2, N, 1, L), self.coef)
        capi.permute(self.coef.view(B * 2, N, L), self.perms, self.out.view(B * 2, N, L), 2, 1, trusted=True)
        return self.out

    # -- hipGraph form --------------------------------------------------------
    def capture(self, fused: bool = False):
        fn = self.run_fused if fused else self.run
        if fused:
            self.prepare_fused()
        stream = torch.cuda.Stream(device=self.device)
        stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(stream):
            fn()  # sizes the plan workspaces outside capture
        torch.cuda.current_stream(self.device).wait_stream(stream)
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream):
            fn()
        self.graph = g
        return g

    def replay(self):
        self.graph.replay()
        return self.out

    def close(self):
        if self.fused is not None:
            self.fused.close()
        self.fwd.close()
        self.inv.close()

    def fused_bytes(self) -> int:
        """HBM bytes the fused form moves per batch at its kernel boundaries:
        read x, write/read the two intermediates, read the key, write out."""
        P_ = 8 * self.n * self.L * self.B
        D = self.dnum
        return P_ + D * P_ + D * P_ + 2 * D * P_ + 2 * P_ + 2 * P_ + 2 * P_

    # -- bookkeeping ------------------------------------------------------------
    def algorithmic_bytes(self) -> int:
        """Irreducible HBM bytes per batch: read x, read evk, write output."""
        return 8 * self.n * self.L * (1 + 2 * self.dnum + 2) * self.B

    def chain_bytes(self) -> int:
        """Bytes the six calls move at their own boundaries (per batch)."""
        P_ = 8 * self.n * self.L * self.B
        D = self.dnum
        gd = P_ + D * P_
        ntt = 2 * D * P_
        mm = D * P_ + 2 * D * P_ + 2 * D * P_
        ams = 2 * D * P_ + 2 * P_
        intt = 4 * P_
        perm = 4 * P_
        return gd + ntt + mm + ams + intt + perm
