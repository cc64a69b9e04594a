"""The serving loop (heal_amd.serving): asynchronous continuous batching of
encrypted linear-layer requests.  A request's result is checked bit-exact
against the oracle: the six-call key-switch chain per rotation, then the
oracle's permute / modsum_ttt / modmul_ttt / modsum_ttt.  It must be the same
when served alone and in a mix of requests of other sizes that join and leave
at other rotation indices, and when it shares a limb-bucket pool with wider
requests."""
from __future__ import annotations

from types import SimpleNamespace

import pytest
import torch

from heal_cases import OracleBackend
from keyswitch_oracle import oracle_keyswitch

pytestmark = pytest.mark.gpu

N, ROT, MAXL, D, W = 4096, 4, 6, 3, 20


@pytest.fixture(scope="module")
def engine(gpu):
    from heal_amd.serving import ServingEngine

    return ServingEngine(ns=(N,), rotations=ROT, max_limbs=MAXL, dnum=D, base_bits=W, max_slots=4, seed=3,
                         limb_buckets=(4, 6))


def _expected(engine, req):
    """acc = sum_r pt_r * rot_r(ct), every step on the oracle."""
    from heal_amd.serving import MAT, request_inputs

    O = OracleBackend()
    ring = engine.rings[req.n]
    L, dev, n, R = req.limbs, ring.device, req.n, ring.R
    q = ring.q[:L].contiguous()
    ct, m = request_inputs(req.seed, n, q, dev)
    ct, m, p = ct.cpu(), m.cpu(), q.cpu()
    prm = SimpleNamespace(p=ring.prm.p[:L], perm=ring.prm.perm, twiddles=ring.prm.twiddles[:L],
                          inv_twiddles=ring.prm.inv_twiddles[:L], m_inv=ring.prm.m_inv[:L])
    rows = torch.arange(n) % MAT
    acc = torch.zeros(2, n, L, dtype=torch.int64)
    for r in range(R):
        if r == 0:
            rot = ct
        else:
            perms = ring.perms[r][:L].contiguous()
            ks = SimpleNamespace(n=n, L=L, dnum=ring.dnum, w=ring.w, prm=prm, x=ct[1].unsqueeze(0),
                                 evk=ring.evk(r)[..., :L].contiguous(), perms=perms)
            k = oracle_keyswitch(ks)
            pc0 = torch.zeros(n, L, dtype=torch.int64)
            assert O.permute(ct[0].clone(), perms.cpu(), pc0, 1, 0) is None
            rot0 = torch.zeros(n, L, dtype=torch.int64)
            assert O.modop(1, k[0].contiguous(), pc0, 0, p, 0, rot0) is None
            rot = torch.stack([rot0, k[1]])
        pt = torch.remainder(m[rows, (rows + r) % MAT].unsqueeze(1), p.unsqueeze(0))  # [n, L]
        prod = torch.zeros(2, n, L, dtype=torch.int64)
        assert O.modop(0, rot.contiguous(), pt.unsqueeze(0).expand(2, n, L), 0, p, 0, prod) is None
        nxt = torch.zeros(2, n, L, dtype=torch.int64)
        assert O.modop(1, acc, prod, 0, p, 0, nxt) is None
        acc = nxt
    return acc


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
           Request(rid=4, n=N, limbs=5, seed=504, arrival=0.003),  # shares the 6-limb pool
           Request(rid=5, n=N, limbs=4, seed=505, arrival=0.006)]
    stats = engine.serve(mix, keep_results=True)
    assert stats["requests"] == 5
    assert torch.equal(mix[2].result.cpu(), exp)
    assert torch.equal(mix[1].result.cpu(), _expected(engine, mix[1]))
    assert torch.equal(mix[3].result.cpu(), _expected(engine, mix[3]))
    # the timeline holds every step once, in device order per pool
    tl = stats["timeline"]
    assert sum(len(t["finished"]) for t in tl) == 5
    for key in {tuple(t["pool"]) for t in tl}:
        ends = [t["end_s"] for t in tl if tuple(t["pool"]) == key]
        assert ends == sorted(ends)


def test_slots_and_hbm_admission(engine):
    from heal_amd.serving import Request

    reqs = [Request(rid=10 + i, n=N, limbs=4, seed=600 + i, arrival=0.0) for i in range(7)]
    stats = engine.serve(reqs)
    assert stats["requests"] == 7 and stats["max_active"] <= engine.max_slots
    assert stats["p50_latency_s"] <= stats["p99_latency_s"]
    for r in reqs:
        assert r.arrival <= r.admitted <= r.finished
    big = Request(rid=99, n=N, limbs=4, seed=1, arrival=0.0)
    saved = engine.hbm_reserve
    try:
        free, _ = torch.cuda.mem_get_info()
        engine.hbm_reserve = free  # nothing fits
        assert not engine._can_admit(big)
    finally:
        engine.hbm_reserve = saved
    assert engine._can_admit(big)


def test_rot_mac_equals_the_five_calls(gpu):
    """hx_rot_mac against permute + modsum_ttt + modmul_ttt + modsum_ttt on the
    device, with signed and unreduced operands (the ops' C++ semantics)."""
    from heal_amd import capi
    from heal_amd import params as P

    B, n, L, M = 3, 1024, 5, 128
    g = torch.Generator().manual_seed(11)
    q = torch.tensor(P.keyswitch_primes(L), dtype=torch.int64)
    def draw(*shape):
        return torch.randint(-(1 << 62), 1 << 62, shape, generator=g, dtype=torch.int64)
    acc, ks, ct = draw(B, 2, n, L), draw(B, 2, n, L), draw(B, 2, n, L)
    acc[0] %= q
    ks[1] %= q
    diags = draw(B, M, M, L)
    diags[2] %= q
    perm = torch.randperm(n, generator=g)
    diag = 37
    dev = {k: v.cuda() for k, v in dict(acc=acc, ks=ks, ct=ct, diags=diags, perm=perm, q=q).items()}
    # the five calls
    rows = torch.arange(n, device="cuda")
    pt = dev["diags"][:, diag, rows % M, :].contiguous()  # [B, n, L]
    pc0 = torch.empty(B, n, L, dtype=torch.int64, device="cuda")
    capi.permute(dev["ct"][:, 0], dev["perm"].view(1, n).expand(L, n), pc0, 2, 1, trusted=True)
    rot = dev["ks"].clone()
    capi.modsum_ttt(dev["ks"][:, 0], pc0, dev["q"], rot[:, 0])
    prod = torch.empty_like(rot)
    capi.modmul_ttt(rot, pt.unsqueeze(1), dev["q"], prod)
    want = torch.empty_like(rot)
    capi.modsum_ttt(dev["acc"], prod, dev["q"], want)
    got = dev["acc"].clone()
    capi.rot_mac(got, dev["ks"], dev["ct"], dev["perm"], dev["diags"], diag, dev["q"])
    assert torch.equal(got, want)
    # diagonal 0 without a key-switch: rot = ct
    prod0 = torch.empty_like(rot)
    pt0 = dev["diags"][:, 0, rows % M, :].contiguous()
    capi.modmul_ttt(dev["ct"], pt0.unsqueeze(1), dev["q"], prod0)
    want0 = torch.empty_like(rot)
    capi.modsum_ttt(dev["acc"], prod0, dev["q"], want0)
    got0 = dev["acc"].clone()
    capi.rot_mac(got0, None, dev["ct"], None, dev["diags"], 0, dev["q"])
    assert torch.equal(got0, want0)
    with pytest.raises(ValueError):
        capi.rot_mac(got, dev["ks"], dev["ct"], None, dev["diags"], diag, dev["q"])


@pytest.mark.parametrize("B,n,L", [(3, 4096, 4), (2, 32768, 16), (2, 65536, 24)])
def test_keyswitch_rmac_equals_keyswitch_then_rot_mac(gpu, B, n, L):
    """hx_keyswitch_exec_rmac (the rotate-multiply-accumulate in the key-switch's
    last kernel) against hx_keyswitch_exec_ex followed by hx_rot_mac, with
    signed and unreduced accumulators and diagonals; shared key, per-call row
    map, two rotations of one plan, at three ring sizes (the last pass's window
    is 6 bits at N = 2^12, 8 at N = 2^16, 7 or 8 at N = 2^15 as tuned)."""
    from heal_amd import capi
    from heal_amd import params as P

    M, D = 128, 3
    prm = P.NttParams(n, P.keyswitch_primes(L), device="cuda")
    fwd = capi.NttPlan(prm.p, prm.perm, prm.twiddles)
    inv = capi.NttPlan(prm.p, prm.perm, prm.inv_twiddles, prm.m_inv, inverse=True)
    ks = capi.KeySwitchPlan(fwd, inv, P.galois_perms(n, 5, L, device="cuda"), D, 60)
    g = torch.Generator(device="cuda")
    g.manual_seed(21)
    q = prm.p
    ct = torch.randint(0, 1 << 62, (B, 2, n, L), generator=g, device="cuda") % q
    evk = torch.randint(0, 1 << 62, (D, 1, 2, n, L), generator=g, device="cuda") % q
    key = ks.prepare_key(evk)
    acc0 = torch.randint(-(1 << 62), 1 << 62, (B, 2, n, L), generator=g, device="cuda")
    acc0[0] %= q
    diags = torch.randint(-(1 << 62), 1 << 62, (B, M, M, L), generator=g, device="cuda")
    diags[1] %= q
    for r, diag in ((3, 3), (5, 133)):
        print(f"[serving] exec_rmac N={n} L={L} rotation {r}", flush=True)
        perm = P.galois_perms(n, pow(5, r, 2 * n), 1, device="cuda")[0].contiguous()
        rmap = ks.row_map(perm)
        out = torch.empty_like(ct)
        ks.exec_ex(ct[:, 1], key, out, key_limbs=L, key_count=1, row_map=rmap)
        want = acc0.clone()
        capi.rot_mac(want, out, ct, perm, diags, diag, q)
        got = acc0.clone()
        ks.exec_rmac(ct[:, 1], key, got, ct, diags, diag, key_limbs=L, key_count=1, row_map=rmap)
        torch.cuda.synchronize()
        assert torch.equal(got, want), r
    with pytest.raises(ValueError):  # acc must not overlap ct
        ks.exec_rmac(ct[:, 1], key, ct, ct, diags, 3, key_limbs=L, key_count=1, row_map=rmap)
    acc = acc0.clone()
    with pytest.raises(ValueError):  # M must be a power of two
        ks.exec_rmac(ct[:, 1], key, acc, ct, diags[:, :96, :96].contiguous(), 3, key_limbs=L, key_count=1,
                     row_map=rmap)
    with pytest.raises(ValueError):  # diagonals are numbered from 0
        ks.exec_rmac(ct[:, 1], key, acc, ct, diags, -1, key_limbs=L, key_count=1, row_map=rmap)
    with pytest.raises(ValueError):  # ct must have acc's strides
        ks.exec_rmac(ct[:, 1], key, acc, ct.transpose(0, 1).contiguous().transpose(0, 1), diags, 3, key_limbs=L,
                     key_count=1, row_map=rmap)
    torch.cuda.synchronize()
    assert torch.equal(acc, acc0)  # refused calls leave acc alone
    ks.close()
    fwd.close()
    inv.close()


def test_n16384_l8_eight_rotations_against_the_oracle(gpu):
    """A mid-size ring (N = 2^14, L = 8, 8 rotations, w = 60): two requests
    served together, each against the oracle."""
    from heal_amd.serving import Request, ServingEngine

    eng = ServingEngine(ns=(1 << 14,), rotations=8, max_limbs=8, dnum=3, base_bits=60, max_slots=4, seed=9,
                        limb_buckets=(8,))
    reqs = [Request(rid=0, n=1 << 14, limbs=8, seed=71, arrival=0.0),
            Request(rid=1, n=1 << 14, limbs=7, seed=72, arrival=0.001)]
    print("[serving] N=16384 L=8 R=8", flush=True)
    stats = eng.serve(reqs, keep_results=True)
    assert stats["requests"] == 2
    for r in reqs:
        print(f"[serving] oracle for request {r.rid}", flush=True)
        assert torch.equal(r.result.cpu(), _expected(eng, r)), r.rid
