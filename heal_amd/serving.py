"""Continuous-batching server for encrypted linear layers on one MI355X.

Workload: a request is a plaintext-matrix x ciphertext-vector product with
diagonal encoding, as a HEAL program would express it with the reference's
operators (key-switch = g_decomposition -> ntt -> modmul -> axis_modsum ->
intt -> permute, then modmul/modsum):

    acc = sum_{r < R}  pt_r (.) rot_r(ct)   (mod q_t, elementwise)
    rot_0(ct) = ct,  rot_r(c0, c1) = (pi_r(c0) + KS_r(c1)[0], KS_r(c1)[1])

where ct = (c0, c1) is [2, N, L] (limb axis last), pi_r is the Galois
permutation for x -> x^(5^r mod 2N), KS_r the fused hybrid key-switch with
evaluation key r (its output is already permuted by pi_r), and pt_r the
diagonal-r encoding of the request's 128 x 128 integer matrix M:
pt_r[j, t] = M[j mod 128, (j + r) mod 128] mod q_t.

Serving model:
* ring degrees N are fixed at start-up; for each N the evaluation keys of all
  R - 1 rotations (prepared once, 24 limbs, shared by every request and every
  limb count L <= 24), the Galois row maps and permutations stay resident;
* requests (N, L, seed, arrival time) are admitted while a slot of their
  (N, L) pool is free and free HBM covers the request's buffers;
* each pool cycles through the rotation indices; at every step all admitted
  requests of a pool run rotation r together: one batched key-switch (shared
  key and row map), one permute, and batched modmul/modsum over the pool's
  slots.  A request admitted at rotation r finishes after R steps of its
  pool.  Modular addition is exact and commutative, so a request's result
  does not depend on when it joined or on the other requests in the pool;
* each pool runs on its own stream (default priority), so pools of
  different (N, L) overlap on the device.

Latency is measured from a request's arrival (Poisson arrivals, seeded) to
the end of the step that completes it.
"""
from __future__ import annotations

import argparse
import json
import math
import time
from dataclasses import dataclass

import torch

from heal_amd import capi
from heal_amd import params as P

MAT = 128  # the plaintext matrix is MAT x MAT


def _gen(seed: int, device) -> torch.Generator:
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    return g


def _uniform(shape, q, g, device):
    """int64 [*shape, L] uniform mod q (q: int64 [L] on device)."""
    r = torch.randint(0, 1 << 62, (*shape, q.numel()), generator=g, device=device, dtype=torch.int64)
    return r % q


@dataclass
class Request:
    rid: int
    n: int
    limbs: int
    seed: int
    arrival: float
    admitted: float | None = None
    finished: float | None = None
    steps: int = 0
    result: torch.Tensor | None = None

    @property
    def bytes(self) -> int:
        # ct, acc, ks out, product tmp: 2NL each; pi(c0) and pt: NL each (int64)
        return 10 * self.n * self.limbs * 8


class _Ring:
    """Everything shared by the requests of one ring degree N."""

    def __init__(self, n: int, rotations: int, dnum: int, base_bits: int, max_limbs: int, seed: int, device):
        self.n, self.R, self.dnum, self.w = n, rotations, dnum, base_bits
        self.device = device
        self.primes = P.keyswitch_primes(max_limbs)
        self.q = torch.tensor(self.primes, dtype=torch.int64, device=device)
        self.seed = seed
        self.prm = P.NttParams(n, self.primes, device=device)  # rows of limb t depend on q_t only
        self.plans = {}  # L -> (fwd, inv, ks)
        self.rows = torch.arange(n, device=device) % MAT
        # Galois permutation of rotation r, the same for every limb: [max_limbs, n] with limb stride 0
        self.perms = [None] * rotations
        for r in range(1, rotations):
            self.perms[r] = P.galois_perms(n, pow(5, r, 2 * n), 1, device=device).expand(max_limbs, n)
        full = self.plan(max_limbs)[2]
        self.row_maps = [None] + [full.row_map(self.perms[r][0]) for r in range(1, rotations)]
        self.keys = [None] + [full.prepare_key(self.evk(r)) for r in range(1, rotations)]

    def evk(self, r: int):
        """Evaluation key of rotation r: [dnum, 1, 2, n, max_limbs] uniform mod q (seeded)."""
        g = _gen(self.seed * 100003 + self.n * 131 + r, self.device)
        return _uniform((self.dnum, 1, 2, self.n), self.q, g, self.device)

    def plan(self, L: int):
        """Forward/inverse NTT plans and the fused key-switch plan of the first L primes."""
        if L not in self.plans:
            prm = self.prm
            fwd = capi.NttPlan(prm.p[:L], prm.perm, prm.twiddles[:L])
            inv = capi.NttPlan(prm.p[:L], prm.perm, prm.inv_twiddles[:L], prm.m_inv[:L], inverse=True)
            ks = capi.KeySwitchPlan(fwd, inv, P.galois_perms(self.n, 5, L, device=self.device), self.dnum, self.w)
            self.plans[L] = (fwd, inv, ks)
        return self.plans[L]


class _Pool:
    """Slots of one (N, L): requests [0, active) run together."""

    def __init__(self, ring: _Ring, L: int, max_slots: int):
        self.ring, self.L, self.max_slots = ring, L, max_slots
        self.n = ring.n
        self.q = ring.q[:L].contiguous()
        self.reqs: list[Request] = []
        self.cursor = 0
        self.cap = 0
        self.stream = torch.cuda.Stream(device=ring.device)
        self.ks = ring.plan(L)[2]
        self._grow(min(4, max_slots))

    def _grow(self, cap: int):
        dev, n, L = self.ring.device, self.n, self.L
        def z(*shape):
            return torch.zeros(*shape, dtype=torch.int64, device=dev)
        old = None if self.cap == 0 else (self.ct, self.acc, self.mats)
        self.ct, self.acc, self.out, self.tmp = z(cap, 2, n, L), z(cap, 2, n, L), z(cap, 2, n, L), z(cap, 2, n, L)
        self.pc0, self.pt = z(cap, n, L), z(cap, n, L)
        self.mats = z(cap, L, MAT, MAT)
        if old is not None:
            k = len(self.reqs)
            self.ct[:k], self.acc[:k], self.mats[:k] = old[0][:k], old[1][:k], old[2][:k]
        self.cap = cap

    def admit(self, req: Request):
        if len(self.reqs) == self.cap:
            self._grow(min(self.max_slots, 2 * self.cap))
        s = len(self.reqs)
        g = _gen(req.seed, self.ring.device)
        self.ct[s] = _uniform((2, self.n), self.q, g, self.ring.device)
        m = torch.randint(-8, 9, (MAT, MAT), generator=g, device=self.ring.device, dtype=torch.int64)
        self.mats[s] = m.unsqueeze(0) % self.q.view(-1, 1, 1)
        self.acc[s].zero_()
        self.reqs.append(req)

    def step(self):
        """Rotation `cursor` for every admitted request of the pool."""
        B, r, ring = len(self.reqs), self.cursor, self.ring
        if B == 0:
            return
        p = self.q
        cols = (ring.rows + r) % MAT
        # pt_r[b, j, t] = M_b[j mod 128, (j + r) mod 128] mod q_t
        self.pt[:B] = self.mats[:B][:, :, ring.rows, cols].transpose(1, 2)
        if r == 0:
            rot = self.ct[:B]
        else:
            self.ks.exec_ex(self.ct[:B, 1], ring.keys[r], self.out[:B], key_limbs=len(ring.primes), key_count=1,
                            row_map=ring.row_maps[r])
            capi.permute(self.ct[:B, 0], ring.perms[r][:self.L], self.pc0[:B], 2, 1, trusted=True)
            capi.modsum_ttt(self.out[:B, 0], self.pc0[:B], p, self.out[:B, 0])
            rot = self.out[:B]
        capi.modmul_ttt(rot, self.pt[:B].unsqueeze(1), p, self.tmp[:B])
        capi.modsum_ttt(self.acc[:B], self.tmp[:B], p, self.acc[:B])
        for q in self.reqs:
            q.steps += 1
        self.cursor = (self.cursor + 1) % ring.R

    def retire(self, now: float, keep_results: bool):
        """Removes finished requests, moving the last active slot into each hole."""
        done = []
        s = 0
        while s < len(self.reqs):
            req = self.reqs[s]
            if req.steps < self.ring.R:
                s += 1
                continue
            req.finished = now
            if keep_results:
                req.result = self.acc[s].clone()
            last = len(self.reqs) - 1
            if s != last:
                self.ct[s], self.acc[s], self.mats[s] = self.ct[last], self.acc[last], self.mats[last]
                self.reqs[s] = self.reqs[last]
            self.reqs.pop()
            done.append(req)
        return done


class ServingEngine:
    def __init__(self, ns=(1 << 14, 1 << 15, 1 << 16), rotations: int = 128, max_limbs: int = 24, dnum: int = 3,
                 base_bits: int = 60, max_slots: int = 16, hbm_reserve: int = 8 << 30, seed: int = 0, device="cuda"):
        self.device = torch.device(device)
        self.rings = {n: _Ring(n, rotations, dnum, base_bits, max_limbs, seed, self.device) for n in ns}
        self.pools: dict[tuple[int, int], _Pool] = {}
        self.max_slots, self.hbm_reserve = max_slots, hbm_reserve
        torch.cuda.synchronize()

    def pool(self, n: int, L: int) -> _Pool:
        if (n, L) not in self.pools:
            self.pools[(n, L)] = _Pool(self.rings[n], L, self.max_slots)
        return self.pools[(n, L)]

    def warm(self, classes):
        """Creates the pools (plans, buffers) of the given (N, L) classes ahead of time."""
        for n, L in classes:
            self.pool(n, L)
        torch.cuda.synchronize()

    def _can_admit(self, req: Request) -> bool:
        pool = self.pool(req.n, req.limbs)
        if len(pool.reqs) >= pool.max_slots:
            return False
        free, _ = torch.cuda.mem_get_info(self.device)
        return free - req.bytes > self.hbm_reserve

    def serve(self, requests: list[Request], keep_results: bool = False) -> dict:
        """Runs until every request has finished; returns latency/throughput figures."""
        pending = sorted(requests, key=lambda r: r.arrival)
        done: list[Request] = []
        t0 = time.perf_counter()
        steps = 0
        max_active = 0
        while len(done) < len(requests):
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now and self._can_admit(pending[0]):
                req = pending.pop(0)
                req.admitted = now
                self.pool(req.n, req.limbs).admit(req)
            active = [p for p in self.pools.values() if p.reqs]
            if not active:
                if pending:  # idle until the next arrival
                    time.sleep(max(0.0, min(pending[0].arrival - now, 0.001)))
                continue
            max_active = max(max_active, sum(len(p.reqs) for p in active))
            main = torch.cuda.current_stream()
            for p in active:
                p.stream.wait_stream(main)
                with torch.cuda.stream(p.stream):
                    p.step()
            for p in active:
                main.wait_stream(p.stream)
            torch.cuda.synchronize()
            steps += 1
            now = time.perf_counter() - t0
            for p in active:
                done += p.retire(now, keep_results)
        wall = time.perf_counter() - t0
        lat = sorted(r.finished - r.arrival for r in done)
        def pct(x):
            return lat[min(len(lat) - 1, max(0, math.ceil(x / 100 * len(lat)) - 1))]
        return {"requests": len(done), "wall_s": wall, "requests_per_s": len(done) / wall,
                "p50_latency_s": pct(50), "p99_latency_s": pct(99), "steps": steps, "max_active": max_active}


def poisson_requests(count: int, rate: float, ns, limbs_range, seed: int) -> list[Request]:
    """Requests with exponential inter-arrival times (mean 1/rate), N and L uniform."""
    g = torch.Generator().manual_seed(seed)
    t, out = 0.0, []
    for i in range(count):
        t += float(-math.log(1.0 - torch.rand(1, generator=g).item()) / rate)
        n = ns[int(torch.randint(0, len(ns), (1,), generator=g))]
        L = int(torch.randint(limbs_range[0], limbs_range[1] + 1, (1,), generator=g))
        out.append(Request(rid=i, n=n, limbs=L, seed=seed * 1000003 + i, arrival=t))
    return out


def main():
    ap = argparse.ArgumentParser(description="serving loop of encrypted linear layers (continuous batching)")
    ap.add_argument("--requests", type=int, default=48)
    ap.add_argument("--rate", type=float, default=20.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--rotations", type=int, default=128)
    ap.add_argument("--max-slots", type=int, default=16)
    ap.add_argument("--seed", type=int, default=7)
    args = ap.parse_args()
    ns, limbs = (1 << 14, 1 << 15, 1 << 16), (8, 24)
    t = time.perf_counter()
    eng = ServingEngine(ns=ns, rotations=args.rotations, max_slots=args.max_slots, seed=args.seed)
    reqs = poisson_requests(args.requests, args.rate, ns, limbs, args.seed)
    eng.warm(sorted({(r.n, r.limbs) for r in reqs}))
    setup = time.perf_counter() - t
    stats = eng.serve(reqs)
    stats.update({"metric": "encrypted linear-layer requests/s (128x128 matrix, diagonal method, "
                            f"{args.rotations} rotations by permute + key-switch)",
                  "rate_req_per_s": args.rate, "setup_s": round(setup, 2), "ns": list(ns), "limbs": list(limbs),
                  "key_switches_per_s": round(args.requests * (args.rotations - 1) / stats["wall_s"], 1)})
    for k in ("wall_s", "requests_per_s", "p50_latency_s", "p99_latency_s"):
        stats[k] = round(stats[k], 4)
    print(json.dumps(stats), flush=True)


if __name__ == "__main__":
    main()
