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
        perms = P.galois_perms(N, galois_elt, L, domain="coeff", device=self.device)  # [L, N]
        dev = self.device
        self.digits = torch.empty((B, N, L, D), dtype=torch.int64, device=dev)
        self.dntt = torch.empty((B, N, D, L), dtype=torch.int64, device=dev)
        # products [D, B, 2, N, L] stored as [B, N, 2, D, L]: modmul_ttt walks its
        # result in memory order, so the digit NTTs it reads ([B, N, D, L]) come
        # in whole lines and serve both output polynomials back to back, and
        # axis_modsum reads the D terms of each output next to each other
        self.prod = torch.empty((B, N, 2, D, L), dtype=torch.int64, device=dev).permute(3, 0, 2, 1, 4)
        self.acc = torch.empty((B, 2, N, L), dtype=torch.int64, device=dev)
        self.coef = torch.empty((B * 2, N, 1, L), dtype=torch.int64, device=dev)
        self.out = torch.empty((B, 2, N, L), dtype=torch.int64, device=dev)
        self.fwd = capi.NttPlan(self.prm.p, self.prm.perm, self.prm.twiddles)
        self.inv = capi.NttPlan(self.prm.p, self.prm.perm, self.prm.inv_twiddles, self.prm.m_inv, inverse=True)
        self.graph = None
        self.fused = None
        self.fused_ok = True  # False: the fused kernels refused these inputs; run_fused runs the chain
        self.key = None
        self.perms = perms

    # -- the Galois map --------------------------------------------------------
    @property
    def perms(self):
        return self._perms

    @perms.setter
    def perms(self, perms):
        """Validates the map once when it is assigned (synchronous; IndexError if
        an entry is out of range) and records whether every limb has the same
        row, so that the chain's permute reads row 0 only.  A map edited in place
        must be assigned again."""
        B, N, L = self.B, self.n, self.L
        capi.permute(self.coef.view(B * 2, N, L), perms, self.out.view(B * 2, N, L), 2, 1)
        self.perms_same = capi.perms_rows_same(perms)
        self._perms = perms
        if self.fused is not None:  # the fused plan holds the row map of the old table
            self.fused.close()
            self.fused = None
            self.key = None
        self.fused_ok = True

    # -- fused form: three kernels, key in device layout ----------------------
    def prepare_fused(self):
        """The fused plan and device key, or None when the fused kernels refuse
        the inputs: keys with negative words (the chain then keeps the signed
        values of the reference's modmul through its INTT), tables outside
        [0, q) or digits too wide for the moduli.  Those run the chain."""
        if self.fused is None and self.fused_ok:
            plan = None
            try:
                plan = capi.KeySwitchPlan(self.fwd, self.inv, self.perms, self.dnum, self.w)
                self.key = plan.prepare_key(self.evk)
            except ValueError as e:
                if "six-call chain" not in str(e):
                    raise
                if plan is not None:
                    plan.close()
                self.fused_ok = False
                return None
            self.fused = plan
        return self.fused

    def run_fused(self, out=None):
        out = self.out if out is None else out
        if self.prepare_fused() is None:
            res = self.run()
            if out is not res:
                out.copy_(res)
            return out
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
        capi.permute(self.coef.view(B * 2, N, L), self.perms, self.out.view(B * 2, N, L), 2, 1, trusted=True,
                     rows_same=self.perms_same)
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


class UnevenKeySwitch:
    """`count` independent key-switches of mixed size: N uniform in `ns`, L
    uniform in [limbs[0], limbs[1]] (seeded); the primes of an instance are the
    first L of the 24-prime key-switch list (q = 1 mod 2^17, so 1 mod 2N for
    every N here); x and the evaluation key uniform mod q.  Instances of the
    same (N, L) share plans.

    run_plain() runs them one key-switch call at a time; group() builds the
    grouped form (one launch per phase and window size for all of them)."""

    def __init__(self, count: int = 64, ns=(1 << 13, 1 << 14, 1 << 15, 1 << 16), limbs=(4, 24), dnum: int = 3,
                 base_bits: int = 60, seed: int = 2, device: str = "cuda"):
        self.device = torch.device(device)
        self.dnum, self.w = dnum, base_bits
        g = torch.Generator().manual_seed(seed)
        self.shapes = [(ns[int(torch.randint(0, len(ns), (1,), generator=g))],
                        int(torch.randint(limbs[0], limbs[1] + 1, (1,), generator=g))) for _ in range(count)]
        self.primes = P.keyswitch_primes(max(L for _, L in self.shapes))
        self.params = {n: P.NttParams(n, self.primes, device=self.device) for n in sorted({n for n, _ in self.shapes})}
        self.plans = {}
        gd = torch.Generator(device=self.device)
        gd.manual_seed(seed)
        q = torch.tensor(self.primes, dtype=torch.int64, device=self.device)
        self.x, self.evk, self.keys, self.outs, self.members = [], [], [], [], []
        for n, L in self.shapes:
            plan = self.plan(n, L)
            x = torch.randint(0, 1 << 62, (1, n, L), generator=gd, device=self.device) % q[:L]
            evk = torch.randint(0, 1 << 62, (dnum, 1, 2, n, L), generator=gd, device=self.device) % q[:L]
            self.x.append(x)
            self.evk.append(evk)
            self.keys.append(plan.prepare_key(evk))
            self.outs.append(torch.empty((1, 2, n, L), dtype=torch.int64, device=self.device))
            self.members.append(plan)

    def plan(self, n: int, L: int):
        if (n, L) not in self.plans:
            prm = self.params[n]
            fwd = capi.NttPlan(prm.p[:L], prm.perm, prm.twiddles[:L])
            inv = capi.NttPlan(prm.p[:L], prm.perm, prm.inv_twiddles[:L], prm.m_inv[:L], inverse=True)
            perms = P.galois_perms(n, 5, L, device=self.device)
            self.plans[(n, L)] = (capi.KeySwitchPlan(fwd, inv, perms, self.dnum, self.w), fwd, inv, perms)
        return self.plans[(n, L)][0]

    def run_plain(self, outs=None):
        outs = self.outs if outs is None else outs
        for plan, x, key, out in zip(self.members, self.x, self.keys, outs):
            plan(x, key, out)
        return outs

    def group(self, outs=None):
        return capi.KeySwitchGroup(self.members, self.x, self.keys, self.outs if outs is None else outs)

    def algorithmic_bytes(self) -> int:
        """Read x and the key, write the output: 8 N L (1 + 2 dnum + 2) per instance."""
        return sum(8 * n * L * (1 + 2 * self.dnum + 2) for n, L in self.shapes)


def _uneven_main():
    import argparse
    import json
    import time

    ap = argparse.ArgumentParser(description="uneven batch of key-switches: plain calls vs one grouped launch")
    ap.add_argument("--count", type=int, default=64)
    ap.add_argument("--reps", type=int, default=20)
    args = ap.parse_args()
    t = time.perf_counter()
    u = UnevenKeySwitch(count=args.count)
    setup = time.perf_counter() - t
    grp = u.group([torch.empty_like(o) for o in u.outs])
    u.run_plain()
    grp()
    torch.cuda.synchronize()
    equal = all(torch.equal(a, b) for a, b in zip(u.outs, grp._keep[3]))
    graph = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        grp()
    torch.cuda.current_stream().wait_stream(s)
    with torch.cuda.graph(graph):
        grp()

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

    t_plain, t_group, t_graph = timed(u.run_plain), timed(grp), timed(graph.replay)
    n = args.count
    print(json.dumps({
        "metric": f"key-switches/s, {n} instances of mixed N in 2^13..2^16 and L in [4, 24] (dnum 3, w 60)",
        "plain_ks_per_s": round(n / t_plain, 1), "grouped_ks_per_s": round(n / t_group, 1),
        "grouped_graph_ks_per_s": round(n / t_graph, 1), "grouped_equal_plain": equal,
        "algorithmic_GBps_graph": round(u.algorithmic_bytes() / t_graph / 1e9, 1),
        "shapes": sorted(u.shapes), "setup_s": round(setup, 2)}), flush=True)


if __name__ == "__main__":
    _uneven_main()
