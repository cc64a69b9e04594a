"""Continuous-batching server for encrypted linear layers on one MI355X.

Workload: a request is a plaintext-matrix x ciphertext-vector product with
diagonal encoding, as a HEAL program would express it with the reference's
operators (key-switch = g_decomposition -> ntt -> modmul -> axis_modsum ->
intt -> permute, then permute/modsum/modmul/modsum):

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
* a request (N, L) joins the pool (N, L') with L' the smallest limb bucket
  >= L (limbs are independent in every step, so the first L limbs of a pool
  of L' limbs are exactly the request's computation; the others hold zeros);
  it is admitted while a slot of its pool is free and free HBM covers the
  request's share;
* each pool cycles through the rotation indices; one step runs rotation r
  for all admitted requests of the pool: one batched key-switch (shared key
  and row map) whose last kernel applies the rotate-multiply-accumulate to the
  accumulators (hx_keyswitch_exec_rmac; rotation 0 is one hx_rot_mac).  A
  request admitted at rotation r finishes after R steps of its pool.  Modular addition is exact and commutative, so a request's
  result does not depend on when it joined or on the other requests;
* every pool runs on its own stream (default priority) and the host never
  waits for the device: it enqueues up to `depth` steps per pool ahead,
  admits new requests into free slots (their set-up is queued on the pool's
  stream, so it lands between steps) and retires finished requests as soon
  as the event of the step that completed them has fired.  A request's
  slot is reused at once: its result copy and the slot compaction are queued
  on the same stream behind the step that finished it.

Latency runs from a request's arrival (Poisson arrivals, seeded) to the
device end time of the step that completes it (events with timing, against
an event recorded at the start of serve()).  serve() also returns a per-step
timeline: pool, batch, rotation, host enqueue time and duration, and device
start/end.
"""
from __future__ import annotations

import argparse
import collections
import gc
import json
import math
import time
from dataclasses import dataclass, field

import torch

from heal_amd import capi
from heal_amd import params as P

MAT = 128  # the plaintext matrix is MAT x MAT
LIMB_BUCKETS = (8, 12, 16, 20, 24)


def _gen(seed: int, device) -> torch.Generator:
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    return g


def _uniform(shape, q, g, device):
    """int64 [*shape, L] uniform mod q (q: int64 [L] on device)."""
    r = torch.randint(0, 1 << 62, (*shape, q.numel()), generator=g, device=device, dtype=torch.int64)
    return r % q


def request_inputs(seed: int, n: int, q, device):
    """The ciphertext [2, n, L] and matrix [MAT, MAT] of a request (seeded; q = its L moduli)."""
    g = _gen(seed, device)
    ct = _uniform((2, n), q, g, device)
    m = torch.randint(-8, 9, (MAT, MAT), generator=g, device=device, dtype=torch.int64)
    return ct, m


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
        # ct and acc: 2NL each (int64), plus the matrix
        return (4 * self.n + MAT * MAT) * self.limbs * 8


@dataclass
class _Step:
    pool: tuple
    r: int
    batch: int
    enqueued: float
    start: torch.cuda.Event
    end: torch.cuda.Event
    finished: list = field(default_factory=list)
    enqueue_s: float = 0.0  # host time spent enqueueing the step (its device interval includes it)


class _Ring:
    """Everything shared by the requests of one ring degree N."""

    def __init__(self, n: int, rotations: int, dnum: int, base_bits: int, max_limbs: int, seed: int, device):
        self.n, self.R, self.dnum, self.w = n, rotations, dnum, base_bits
        self.device = device
        self.max_limbs = max_limbs
        self.primes = P.keyswitch_primes(max_limbs)
        self.q = torch.tensor(self.primes, dtype=torch.int64, device=device)
        self.seed = seed
        self.prm = P.NttParams(n, self.primes, device=device)  # rows of limb t depend on q_t only
        self.plans = {}  # L -> (fwd, inv, ks)
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
    """Slots of one (N, L'): requests [0, len(reqs)) run together, on the pool's own stream."""

    def __init__(self, ring: _Ring, L: int, max_slots: int):
        self.ring, self.L, self.max_slots = ring, L, max_slots
        self.n = ring.n
        self.key = (ring.n, L)
        self.q = ring.q[:L].contiguous()
        self.reqs: list[Request] = []
        self.cursor = 0
        self.inflight: collections.deque[_Step] = collections.deque()
        self.stream = torch.cuda.Stream(device=ring.device)
        self.ks = ring.plan(L)[2]
        dev, n = ring.device, self.n
        with torch.cuda.stream(self.stream):
            self.ct = torch.zeros(max_slots, 2, n, L, dtype=torch.int64, device=dev)
            self.acc = torch.zeros_like(self.ct)
            # the matrices by diagonals: diags[s, r, i, t] = M_s[i, (i + r) mod MAT] mod q_t
            self.diags = torch.zeros(max_slots, MAT, MAT, L, dtype=torch.int64, device=dev)

    def admit(self, req: Request):
        """Queues the request's set-up on the pool's stream (behind the steps already queued)."""
        s, Lr, dev = len(self.reqs), req.limbs, self.ring.device
        with torch.cuda.stream(self.stream):
            ct, m = request_inputs(req.seed, self.n, self.q[:Lr], dev)
            self.ct[s].zero_()
            self.ct[s, :, :, :Lr] = ct
            i = torch.arange(MAT, device=dev)
            dg = m[i.view(1, -1), (i.view(1, -1) + i.view(-1, 1)) % MAT]  # [r, i]
            self.diags[s].zero_()
            self.diags[s, :, :, :Lr] = dg.unsqueeze(-1) % self.q[:Lr]
            self.acc[s].zero_()
        self.reqs.append(req)

    def step(self, now: float, keep_results: bool) -> _Step:
        """Queues rotation `cursor` for every admitted request, then the retirement
        of the requests it completes."""
        B, r, ring = len(self.reqs), self.cursor, self.ring
        h0 = time.perf_counter()
        rec = _Step(self.key, r, B, now, torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        with torch.cuda.stream(self.stream):
            rec.start.record()
            if r == 0:
                capi.rot_mac(self.acc[:B], None, self.ct[:B], None, self.diags[:B], 0, self.q)
            else:
                # the key-switch with the rotate-multiply-accumulate in its last kernel
                self.ks.exec_rmac(self.ct[:B, 1], ring.keys[r], self.acc[:B], self.ct[:B], self.diags[:B], r,
                                  key_limbs=ring.max_limbs, key_count=1, row_map=ring.row_maps[r])
            rec.end.record()
            rec.enqueue_s = time.perf_counter() - h0
            for q in self.reqs:
                q.steps += 1
            rec.finished = self._retire(keep_results)
        self.cursor = (r + 1) % ring.R
        self.inflight.append(rec)
        return rec

    def _retire(self, keep_results: bool):
        """Removes the requests that have run R steps, moving the last slot into each hole
        (queued on the current stream, behind the step)."""
        done = []
        s = 0
        while s < len(self.reqs):
            req = self.reqs[s]
            if req.steps < self.ring.R:
                s += 1
                continue
            if keep_results:
                req.result = self.acc[s, :, :, :req.limbs].clone()
            last = len(self.reqs) - 1
            if s != last:
                self.ct[s].copy_(self.ct[last])
                self.acc[s].copy_(self.acc[last])
                self.diags[s].copy_(self.diags[last])
                self.reqs[s] = self.reqs[last]
            self.reqs.pop()
            done.append(req)
        return done


class ServingEngine:
    def __init__(self, ns=(1 << 14, 1 << 15, 1 << 16), rotations: int = 128, max_limbs: int = 24, dnum: int = 3,
                 base_bits: int = 60, max_slots: int = 16, hbm_reserve: int = 8 << 30, seed: int = 0,
                 limb_buckets=LIMB_BUCKETS, depth: int = 2, device="cuda"):
        self.device = torch.device(device)
        self.rings = {n: _Ring(n, rotations, dnum, base_bits, max_limbs, seed, self.device) for n in ns}
        self.pools: dict[tuple[int, int], _Pool] = {}
        self.max_slots, self.hbm_reserve = max_slots, hbm_reserve
        self.buckets = tuple(sorted(b for b in (limb_buckets or ()) if b <= max_limbs))
        self.max_limbs = max_limbs
        self.depth = depth
        torch.cuda.synchronize()

    def bucket(self, L: int) -> int:
        for b in self.buckets:
            if b >= L:
                return b
        return L

    def pool(self, n: int, L: int) -> _Pool:
        key = (n, self.bucket(L))
        if key not in self.pools:
            self.pools[key] = _Pool(self.rings[n], key[1], self.max_slots)
        return self.pools[key]

    def warm(self, classes):
        """Creates the pools (plans, buffers) of the given (N, L) classes ahead of time
        and runs one step of each kind on every pool's stream with every slot
        filled (first-use set-up: stream workspaces at their largest, generators,
        allocator blocks), then leaves the pools empty."""
        for n, L in classes:
            self.pool(n, L)
        for p in self.pools.values():
            if p.reqs or p.inflight:
                continue
            for s in range(p.max_slots):
                p.admit(Request(rid=-1 - s, n=p.n, limbs=p.L, seed=s, arrival=0.0))
            for _ in range(min(2, p.ring.R)):
                p.step(0.0, False)
            p.reqs.clear()
            p.inflight.clear()
            p.cursor = 0
        torch.cuda.synchronize()

    def _can_admit(self, req: Request) -> bool:
        if req.limbs < 1 or req.limbs > self.max_limbs or req.n not in self.rings:
            raise ValueError(f"request {req.rid}: no ring for N={req.n}, L={req.limbs}")
        pool = self.pool(req.n, req.limbs)
        if len(pool.reqs) >= pool.max_slots:
            return False
        free, _ = torch.cuda.mem_get_info(self.device)
        return free - req.bytes > self.hbm_reserve

    def serve(self, requests: list[Request], keep_results: bool = False) -> dict:
        """Runs until every request has finished; returns latency/throughput figures and the timeline."""
        pending = sorted(requests, key=lambda r: r.arrival)
        done: list[Request] = []
        timeline = []
        torch.cuda.synchronize()
        # the collector's pauses would land inside steps being enqueued (a step's
        # device interval runs from its first to its last enqueued event): it is
        # off while serving, and the objects alive now are left out of later passes
        gc_was_on = gc.isenabled()
        gc.collect()
        gc.freeze()
        gc.disable()
        try:
            return self._serve(pending, requests, keep_results, done, timeline)
        finally:
            if gc_was_on:
                gc.enable()
            gc.unfreeze()

    def _serve(self, pending, requests, keep_results, done, timeline) -> dict:
        origin = torch.cuda.Event(enable_timing=True)
        origin.record()
        origin.synchronize()
        t0 = time.perf_counter()
        steps = 0
        max_active = 0
        while len(done) < len(requests):
            now = time.perf_counter() - t0
            # admission: every arrived request whose pool has a free slot, in arrival order
            i = 0
            while i < len(pending) and pending[i].arrival <= now:
                if self._can_admit(pending[i]):
                    req = pending.pop(i)
                    req.admitted = now
                    self.pool(req.n, req.limbs).admit(req)
                else:
                    i += 1
            # enqueue up to `depth` steps per pool
            active = 0
            for p in self.pools.values():
                while p.reqs and len(p.inflight) < self.depth:
                    p.step(time.perf_counter() - t0, keep_results)
                    steps += 1
                active += len(p.reqs)
            max_active = max(max_active, active)
            # completions, in stream order per pool
            progressed = False
            for p in self.pools.values():
                while p.inflight and p.inflight[0].end.query():
                    rec = p.inflight.popleft()
                    start = origin.elapsed_time(rec.start) / 1e3
                    end = origin.elapsed_time(rec.end) / 1e3
                    for req in rec.finished:
                        req.finished = end
                        done.append(req)
                    timeline.append({"pool": list(rec.pool), "r": rec.r, "batch": rec.batch,
                                     "enqueued_s": round(rec.enqueued, 6), "enqueue_ms": round(rec.enqueue_s * 1e3, 3),
                                     "start_s": round(start, 6), "end_s": round(end, 6),
                                     "finished": [q.rid for q in rec.finished]})
                    progressed = True
            if not progressed:
                waiting = any(p.inflight for p in self.pools.values())
                if waiting or pending:
                    time.sleep(0.0002)
        torch.cuda.synchronize()
        wall = max((r.finished for r in done), default=0.0)
        lat = sorted(r.finished - r.arrival for r in done)
        queue = sorted(r.admitted - r.arrival for r in done)

        def pct(v, x):
            return v[min(len(v) - 1, max(0, math.ceil(x / 100 * len(v)) - 1))] if v else 0.0

        # device busy time: union of the step intervals over all pools
        busy, cur_s, cur_e = 0.0, None, None
        for s, e in sorted((t["start_s"], t["end_s"]) for t in timeline):
            if cur_e is None or s > cur_e:
                if cur_e is not None:
                    busy += cur_e - cur_s
                cur_s, cur_e = s, e
            else:
                cur_e = max(cur_e, e)
        if cur_e is not None:
            busy += cur_e - cur_s
        return {"requests": len(done), "wall_s": wall, "requests_per_s": len(done) / wall if wall else 0.0,
                "p50_latency_s": pct(lat, 50), "p99_latency_s": pct(lat, 99),
                "p50_queue_s": pct(queue, 50), "p99_queue_s": pct(queue, 99),
                "steps": steps, "max_active": max_active, "device_busy_frac": busy / wall if wall else 0.0,
                "timeline": timeline}


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


def summarize_timeline(timeline, done_ids=None) -> dict:
    """Per-pool step counts, mean batch and mean device step time."""
    per = {}
    for t in timeline:
        k = f"{t['pool'][0]}x{t['pool'][1]}"
        e = per.setdefault(k, {"steps": 0, "batch": 0, "device_s": 0.0})
        e["steps"] += 1
        e["batch"] += t["batch"]
        e["device_s"] += t["end_s"] - t["start_s"]
    return {k: {"steps": v["steps"], "mean_batch": round(v["batch"] / v["steps"], 2),
                "mean_step_ms": round(1e3 * v["device_s"] / v["steps"], 3)} for k, v in sorted(per.items())}


def main():
    ap = argparse.ArgumentParser(description="serving loop of encrypted linear layers (continuous batching)")
    ap.add_argument("--requests", type=int, default=48)
    ap.add_argument("--rates", default="20,sat", help="comma list of Poisson rates (req/s); 'sat' = all at t=0")
    ap.add_argument("--sat-requests", type=int, default=240, help="requests of the saturation run")
    ap.add_argument("--rotations", type=int, default=128)
    ap.add_argument("--max-slots", type=int, default=16)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--timeline", default=None, help="write the per-step timelines to this JSON file")
    ap.add_argument("--ns", default="16384,32768,65536", help="ring degrees of the requests (uniform)")
    ap.add_argument("--limbs", default="8,24", help="limb range of the requests (uniform, inclusive)")
    args = ap.parse_args()
    ns = tuple(int(x) for x in args.ns.split(","))
    limbs = tuple(int(x) for x in args.limbs.split(","))
    t = time.perf_counter()
    eng = ServingEngine(ns=ns, rotations=args.rotations, max_slots=args.max_slots, seed=args.seed)
    eng.warm([(n, eng.bucket(L)) for n in ns for L in range(limbs[0], limbs[1] + 1)])
    setup = time.perf_counter() - t
    timelines = {}
    for rate in args.rates.split(","):
        sat = rate == "sat"
        count = args.sat_requests if sat else args.requests
        reqs = poisson_requests(count, 1e9 if sat else float(rate), ns, limbs, args.seed)
        stats = eng.serve(reqs)
        timelines[rate] = stats.pop("timeline")
        stats.update({"metric": "encrypted linear-layer requests/s (128x128 matrix, diagonal method, "
                                f"{args.rotations} rotations by permute + key-switch)",
                      "rate_req_per_s": "saturation" if sat else float(rate), "setup_s": round(setup, 2),
                      "ns": list(ns), "limbs": list(limbs), "limb_buckets": list(eng.buckets),
                      "depth": eng.depth,
                      "key_switches_per_s": round(count * (args.rotations - 1) / stats["wall_s"], 1),
                      "pools": summarize_timeline(timelines[rate])})
        for k in ("wall_s", "requests_per_s", "p50_latency_s", "p99_latency_s", "p50_queue_s", "p99_queue_s",
                  "device_busy_frac"):
            stats[k] = round(stats[k], 4)
        print(json.dumps(stats), flush=True)
    if args.timeline:
        with open(args.timeline, "w") as f:
            json.dump(timelines, f)


if __name__ == "__main__":
    main()
