"""The serving loop (heal_amd.serving): continuous batching of encrypted
linear-layer requests.  A request's result is checked bit-exact against the
oracle's six-call key-switch chain (per rotation) plus exact modular
arithmetic, and must be the same when served alone and in a mix of requests
of other sizes that join and leave at other rotation indices."""
from __future__ import annotations

from types import SimpleNamespace

import pytest
import torch

from keyswitch_oracle import oracle_keyswitch

pytestmark = pytest.mark.gpu

N, ROT, MAXL, D, W = 4096, 4, 6, 3, 20


@pytest.fixture(scope="module")
def engine(gpu):
    from heal_amd.serving import ServingEngine

    return ServingEngine(ns=(N,), rotations=ROT, max_limbs=MAXL, dnum=D, base_bits=W, max_slots=4, seed=3)


def _expected(engine, req):
    """acc = sum_r pt_r * rot_r(ct) from the oracle key-switch chain, Python integers."""
    from heal_amd.serving import MAT, _gen, _uniform

    ring = engine.rings[req.n]
    L, dev = req.limbs, ring.device
    q = ring.q[:L].contiguous()
    g = _gen(req.seed, dev)
    ct = _uniform((2, req.n), q, g, dev).cpu()
    m = torch.randint(-8, 9, (MAT, MAT), generator=g, device=dev, dtype=torch.int64).cpu()
    primes = ring.primes[:L]
    prm = SimpleNamespace(p=ring.prm.p[:L], perm=ring.prm.perm, twiddles=ring.prm.twiddles[:L],
                          inv_twiddles=ring.prm.inv_twiddles[:L], m_inv=ring.prm.m_inv[:L])
    acc = [[[0] * L for _ in range(req.n)] for _ in range(2)]
    c0 = ct[0].tolist()
    for r in range(ROT):
        if r == 0:
            rot = ct.tolist()
        else:
            perms = ring.perms[r][:L].contiguous()
            ks = SimpleNamespace(n=req.n, L=L, dnum=D, w=W, prm=prm, x=ct[1].unsqueeze(0),
                                 evk=ring.evk(r)[..., :L].contiguous(), perms=perms)
            k = oracle_keyswitch(ks).tolist()
            pr = perms[0].tolist()
            rot = [[[(c0[pr[j]][t] + k[0][j][t]) % primes[t] for t in range(L)] for j in range(req.n)], k[1]]
        for c in range(2):
            for j in range(req.n):
                row = m[j % MAT, (j + r) % MAT].item()
                for t in range(L):
                    acc[c][j][t] = (acc[c][j][t] + (row % primes[t]) * rot[c][j][t]) % primes[t]
    return torch.tensor(acc, dtype=torch.int64)


def test_alone_and_in_the_mix_agree_with_the_oracle(engine):
    from heal_amd.serving import Request

    print("[serving] alone", flush=True)
    alone = Request(rid=0, n=N, limbs=4, seed=501, arrival=0.0)
    engine.serve([alone], keep_results=True)
    exp = _expected(engine, alone)
    assert torch.equal(alone.result.cpu(), exp)
    print("[serving] mix", flush=True)
    mix = [Request(rid=1, n=N, limbs=4, seed=502, arrival=0.0),
           Request(rid=2, n=N, limbs=6, seed=503, arrival=0.0),
           Request(rid=3, n=N, limbs=4, seed=501, arrival=0.003),  # the same request as `alone`
           Request(rid=4, n=N, limbs=5, seed=504, arrival=0.003),
           Request(rid=5, n=N, limbs=4, seed=505, arrival=0.006)]
    stats = engine.serve(mix, keep_results=True)
    assert stats["requests"] == 5
    assert torch.equal(mix[2].result.cpu(), exp)
    assert torch.equal(mix[1].result.cpu(), _expected(engine, mix[1]))


def test_slots_and_hbm_admission(engine):
    from heal_amd.serving import Request

    reqs = [Request(rid=10 + i, n=N, limbs=4, seed=600 + i, arrival=0.0) for i in range(7)]
    stats = engine.serve(reqs)
    assert stats["requests"] == 7 and stats["max_active"] <= engine.max_slots
    assert stats["p50_latency_s"] <= stats["p99_latency_s"]
    big = Request(rid=99, n=N, limbs=4, seed=1, arrival=0.0)
    saved = engine.hbm_reserve
    try:
        free, _ = torch.cuda.mem_get_info()
        engine.hbm_reserve = free  # nothing fits
        assert not engine._can_admit(big)
    finally:
        engine.hbm_reserve = saved
    assert engine._can_admit(big)
