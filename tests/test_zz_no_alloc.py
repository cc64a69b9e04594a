"""hx_ntt_plan_exec_noalloc (the NTT with the library's fixed device scratch,
no allocation) against hx_ntt_plan_exec through the C ABI: two calls in a
row, the second with a smaller input, each equal bit for bit; columns with
negative inputs exercise the part of the scratch the reference-semantics
kernel uses; a transform larger than the scratch is refused."""
from __future__ import annotations

import pytest
import torch

from heal_amd import params as P

pytestmark = pytest.mark.gpu


def _input(primes, l, m, r, g, signed):
    a = torch.stack([torch.randint(0, int(q), (l, m, r), generator=g) for q in primes], dim=-1)
    if signed:  # some negative entries: flagged columns
        a[0, ::5, 0] -= torch.tensor(primes)
    return a.cuda()


def test_noalloc_equals_plan_exec_twice_then_smaller(gpu):
    from heal_amd import capi

    g = torch.Generator().manual_seed(17)
    runs = [(65536, P.keyswitch_primes(12), 2, 2), (4096, P.ntt_primes(4096, [50, 31, 20]), 1, 3)]
    for i, (m, primes, l, r) in enumerate(runs):
        print(f"[noalloc] call {i}: m={m} k={len(primes)} l={l} r={r}", flush=True)
        prm = P.NttParams(m, primes).to("cuda")
        fwd = capi.NttPlan(prm.p, prm.perm, prm.twiddles)
        inv = capi.NttPlan(prm.p, prm.perm, prm.inv_twiddles, prm.m_inv, inverse=True)
        a = _input(primes, l, m, r, g, signed=True)
        want = torch.empty_like(a)
        got = torch.full_like(a, -7)
        fwd(a, want)
        fwd.exec_noalloc(a, got)
        torch.cuda.synchronize()
        assert torch.equal(got, want), i
        back_want = torch.empty_like(a)
        back_got = torch.full_like(a, -7)
        inv(want, back_want)
        inv.exec_noalloc(want, back_got)
        torch.cuda.synchronize()
        assert torch.equal(back_got, back_want), i
        fwd.close()
        inv.close()


def test_noalloc_refuses_a_transform_larger_than_its_scratch(gpu):
    from heal_amd import capi

    m = 65536
    primes = P.keyswitch_primes(24)
    prm = P.NttParams(m, primes).to("cuda")
    fwd = capi.NttPlan(prm.p, prm.perm, prm.twiddles)
    a = torch.zeros((8, m, 1, 24), dtype=torch.int64, device="cuda")  # 192 columns: 96 MiB of workspace
    with pytest.raises(ValueError, match="workspace"):
        fwd.exec_noalloc(a, torch.empty_like(a))
    fwd.close()
