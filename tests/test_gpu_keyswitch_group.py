"""Grouped key-switches of mixed (N, L) (hx_keyswitch_group_*): every member's
output equals its own hx_keyswitch_exec and the oracle's six-call chain, also
when replayed from a hipGraph and when members share a plan."""
from __future__ import annotations

from types import SimpleNamespace

import pytest
import torch

from keyswitch_oracle import oracle_keyswitch

pytestmark = pytest.mark.gpu


def _oracle(u, i):
    n, L = u.shapes[i]
    prm = u.params[n]
    plan, fwd, inv, perms = u.plans[(n, L)]
    ks = SimpleNamespace(n=n, L=L, dnum=u.dnum, w=u.w, x=u.x[i], evk=u.evk[i], perms=perms,
                         prm=SimpleNamespace(p=prm.p[:L], perm=prm.perm, twiddles=prm.twiddles[:L],
                                             inv_twiddles=prm.inv_twiddles[:L], m_inv=prm.m_inv[:L]))
    return oracle_keyswitch(ks)


@pytest.mark.parametrize("w", [60, 20])
def test_group_matches_plain_and_oracle(gpu, w):
    from heal_amd.keyswitch import UnevenKeySwitch

    print(f"[group] small w={w}", flush=True)
    u = UnevenKeySwitch(count=7, ns=(4096, 8192), limbs=(2, 5), base_bits=w, seed=11 + w)
    outs = [torch.full_like(o, -1) for o in u.outs]
    grp = u.group(outs)
    grp()
    u.run_plain()
    torch.cuda.synchronize()
    for i in range(7):
        assert torch.equal(outs[i], u.outs[i]), (i, u.shapes[i])
    for i in (0, 3, 6):
        assert torch.equal(outs[i][0].cpu(), _oracle(u, i)), (i, u.shapes[i])
    grp.close()


def test_group_shared_plans_and_graph_replay(gpu):
    from heal_amd.keyswitch import UnevenKeySwitch

    print("[group] graph", flush=True)
    u = UnevenKeySwitch(count=10, ns=(4096, 16384), limbs=(3, 4), seed=5)
    assert len(set(u.shapes)) < len(u.shapes)  # some members share a plan
    outs = [torch.empty_like(o) for o in u.outs]
    grp = u.group(outs)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        grp()
    torch.cuda.current_stream().wait_stream(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        grp()
    u.run_plain()
    for rep in range(3):
        for o in outs:
            o.fill_(-7)
        g.replay()
        torch.cuda.synchronize()
        for i in range(len(outs)):
            assert torch.equal(outs[i], u.outs[i]), (rep, i, u.shapes[i])


def test_group_full_size_member(gpu):
    """One N = 2^16, L = 24 member next to small ones."""
    from heal_amd.keyswitch import UnevenKeySwitch

    print("[group] full size", flush=True)
    u = UnevenKeySwitch(count=3, ns=(1 << 16,), limbs=(24, 24), seed=9)
    v = UnevenKeySwitch(count=3, ns=(1 << 13,), limbs=(4, 6), seed=10)
    from heal_amd import capi

    outs = [torch.empty_like(o) for o in u.outs + v.outs]
    grp = capi.KeySwitchGroup(u.members + v.members, u.x + v.x, u.keys + v.keys, outs)
    grp()
    u.run_plain()
    v.run_plain()
    torch.cuda.synchronize()
    for a, b in zip(outs, u.outs + v.outs):
        assert torch.equal(a, b)


def test_group_errors(gpu):
    from heal_amd import capi
    from heal_amd.keyswitch import UnevenKeySwitch

    u = UnevenKeySwitch(count=2, ns=(4096,), limbs=(3, 3), seed=1)
    with pytest.raises(ValueError):
        capi.KeySwitchGroup(u.members, u.x, u.keys, u.outs[:1])
    bad = [torch.empty((1, 2, 4096, 2), dtype=torch.int64, device="cuda"), u.outs[1]]
    with pytest.raises(ValueError):
        capi.KeySwitchGroup(u.members, u.x, u.keys, bad)


@pytest.mark.slow
def test_group_uneven_batch_of_64(gpu):
    """The uneven batch of 64 key-switches (seed 2: N in {2^13..2^16}, L in [4, 24],
    dnum 3, w 60) in one launch per phase: every member equals its own fused
    call, eager and replayed from a hipGraph; sampled members (one of each N)
    equal the oracle's six-call chain."""
    from heal_amd.keyswitch import UnevenKeySwitch

    print("[group] 64 mixed", flush=True)
    u = UnevenKeySwitch(count=64, seed=2)
    outs = [torch.full_like(o, -1) for o in u.outs]
    grp = u.group(outs)
    grp()
    u.run_plain()
    torch.cuda.synchronize()
    for i in range(64):
        assert torch.equal(outs[i], u.outs[i]), (i, u.shapes[i])
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        grp()
    torch.cuda.current_stream().wait_stream(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        grp()
    for rep in range(2):
        for o in outs:
            o.fill_(-7)
        g.replay()
        torch.cuda.synchronize()
        for i in range(64):
            assert torch.equal(outs[i], u.outs[i]), (rep, i, u.shapes[i])
    first = {}
    for i, (n, L) in enumerate(u.shapes):
        first.setdefault(n, i)
    for n, i in sorted(first.items()):
        print(f"[group] oracle member {i}: N={n} L={u.shapes[i][1]}", flush=True)
        assert torch.equal(outs[i][0].cpu(), _oracle(u, i)), (i, u.shapes[i])
    grp.close()
