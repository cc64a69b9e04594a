"""The hybrid key-switch chain on the GPU against the oracle's chain, bit for
bit, in plain (B = 1), batched (B instances per launch) and hipGraph forms;
plus size-independent properties at the full N = 2^16, L = 24 shape."""
from __future__ import annotations

import pytest
import torch

from keyswitch_oracle import oracle_keyswitch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,limbs,w,batch", [(1024, 4, 20, 1), (4096, 6, 60, 3), (8192, 5, 20, 2)])
def test_keyswitch_small_vs_oracle(gpu, n, limbs, w, batch):
    from heal_amd.keyswitch import KeySwitchBatch

    print(f"[keyswitch] n={n} L={limbs} w={w} B={batch}", flush=True)
    ks = KeySwitchBatch(n=n, limbs=limbs, base_bits=w, batch=batch, seed=n + w)
    ks.run()
    torch.cuda.synchronize()
    for b in range(batch):
        assert torch.equal(ks.out[b].cpu(), oracle_keyswitch(ks, b)), b
    ks.close()


def test_keyswitch_graph_replays_match_eager(gpu):
    from heal_amd.keyswitch import KeySwitchBatch

    ks = KeySwitchBatch(n=4096, limbs=8, base_bits=20, batch=2, seed=3)
    ks.run()
    torch.cuda.synchronize()
    want = ks.out.clone()
    ks.capture()
    for i in range(3):
        ks.out.zero_()
        torch.cuda.synchronize()
        ks.replay()
        torch.cuda.synchronize()
        assert torch.equal(ks.out, want), i
        # eager work between replays
        ks.run()
        torch.cuda.synchronize()
        assert torch.equal(ks.out, want), i
    ks.close()


@pytest.mark.parametrize("n,limbs,dnum,w,batch", [(4096, 3, 3, 20, 2), (8192, 5, 2, 30, 1), (16384, 4, 4, 15, 2),
                                                   (32768, 2, 1, 60, 1), (65536, 3, 3, 60, 1)])
def test_fused_keyswitch_vs_oracle(gpu, n, limbs, dnum, w, batch):
    from heal_amd import capi
    from heal_amd.keyswitch import KeySwitchBatch

    print(f"[fused] n={n} L={limbs} D={dnum} w={w} B={batch}", flush=True)
    ks = KeySwitchBatch(n=n, limbs=limbs, dnum=dnum, base_bits=w, batch=batch, seed=n + dnum)
    if n == 8192:
        # unreduced key entries and unreduced / negative x exercise the chain's C++ integer semantics
        # (g_decomposition's unsigned digits, modmul's reduction of unreduced operands)
        ks.evk[:, :, :, ::7] += 1 << 62
        ks.x[:, ::5] = ks.x[:, ::5] * 3 + (1 << 61)
        ks.x[:, 1::9] = -ks.x[:, 1::9]
    capi.status_poll()
    out = torch.zeros_like(ks.out)
    ks.run_fused(out)
    st = capi.status_poll()
    for b in range(batch):
        assert torch.equal(out[b].cpu(), oracle_keyswitch(ks, b)), b
    # the g_decomposition overflow count matches the per-call op
    ks.run()
    st2 = capi.status_poll()
    assert st["gdecomp_overflow"] == st2["gdecomp_overflow"]
    assert torch.equal(out, ks.out)
    ks.close()


def test_fused_keyswitch_negative_key_matches_oracle(gpu):
    """Negative key entries give signed products and digit sums (C++
    truncation), and the chain's INTT keeps the reference's signed
    representatives: the fused plan refuses such a key, the fused call then
    runs the chain, and both equal the oracle's chain."""
    from heal_amd.keyswitch import KeySwitchBatch

    ks = KeySwitchBatch(n=4096, limbs=3, dnum=3, base_bits=20, batch=2, seed=77)
    ks.evk[:, :, :, ::3] = -ks.evk[:, :, :, ::3]
    ks.run()
    torch.cuda.synchronize()
    want = ks.out.clone()
    for b in range(2):
        assert torch.equal(want[b].cpu(), oracle_keyswitch(ks, b)), b
    out = torch.zeros_like(want)
    ks.run_fused(out)
    torch.cuda.synchronize()
    assert not ks.fused_ok
    assert torch.equal(out, want)
    ks.close()


def test_fused_keyswitch_per_limb_permutations(gpu):
    """A different permutation for every limb (the per-limb row scatter in the last
    kernel) and the same one for every limb (one shared row map) both equal the
    six-call chain, whose permute applies perms [L, N] row by row."""
    from heal_amd.keyswitch import KeySwitchBatch

    for shared in (False, True):
        print(f"[fused perms] shared={shared}", flush=True)
        ks = KeySwitchBatch(n=4096, limbs=5, dnum=2, base_bits=30, batch=2, seed=91)
        g = torch.Generator().manual_seed(5)
        rows = [torch.randperm(ks.n, generator=g) for _ in range(1 if shared else ks.L)]
        ks.perms = torch.stack([rows[0 if shared else t] for t in range(ks.L)]).to(ks.perms)
        ks.run()
        torch.cuda.synchronize()
        want = ks.out.clone()
        out = torch.zeros_like(want)
        ks.run_fused(out)
        torch.cuda.synchronize()
        assert torch.equal(out, want), shared
        ks.close()


def test_fused_keyswitch_mixed_prime_classes(gpu):
    """Limbs of every butterfly class in one key-switch: 30/31-bit primes (32-bit
    butterflies, Möller–Granlund inner product), 55/60-bit primes (lazy 64-bit
    butterflies, key in Montgomery form) and 61/62-bit primes (fully reduced
    butterflies, Möller–Granlund inner product).  The fused form equals the
    six-call chain, which equals the oracle's chain."""
    from heal_amd import params as P
    from heal_amd.keyswitch import KeySwitchBatch

    primes = P.ntt_primes(4096, [30, 61, 55, 62, 31, 60])
    print(f"[fused classes] primes={primes}", flush=True)
    ks = KeySwitchBatch(n=4096, limbs=len(primes), dnum=3, base_bits=24, batch=2, seed=123, primes=primes)
    ks.run()
    torch.cuda.synchronize()
    want = ks.out.clone()
    out = torch.zeros_like(want)
    ks.run_fused(out)
    torch.cuda.synchronize()
    assert torch.equal(out, want)
    for b in range(2):
        assert torch.equal(want[b].cpu(), oracle_keyswitch(ks, b))
    ks.close()


def test_fused_keyswitch_full_size_matches_plain_chain(gpu):
    from heal_amd.keyswitch import KeySwitchBatch

    ks = KeySwitchBatch(n=65536, limbs=24, dnum=3, base_bits=20, batch=2, seed=5)
    ks.run()
    torch.cuda.synchronize()
    want = ks.out.clone()
    out = torch.zeros_like(ks.out)
    ks.run_fused(out)
    torch.cuda.synchronize()
    assert torch.equal(out, want)
    # graph replays of the fused form, with eager work in between
    ks.capture(fused=True)
    for i in range(3):
        ks.out.zero_()
        ks.replay()
        torch.cuda.synchronize()
        assert torch.equal(ks.out, want), i
        ks.run_fused(out)
        torch.cuda.synchronize()
        assert torch.equal(out, want), i
    ks.close()


@pytest.mark.parametrize("n,limbs", [(4096, 4), (65536, 24)])
def test_fused_keyswitch_extreme_residues(gpu, n, limbs):
    """Every word of x and of the evaluation key q - 1 (w = 20: all three digits
    live), so that the fused kernels' lazy values reach their bounds (K1's and
    K2's forward stages below 16q, the reduction before the key product, the
    inverse stages below 4q).  Fused equals the six-call chain; at N = 4096
    instance 0 equals the oracle's chain."""
    from heal_amd.keyswitch import KeySwitchBatch

    print(f"[fused extreme] n={n} L={limbs}", flush=True)
    ks = KeySwitchBatch(n=n, limbs=limbs, dnum=3, base_bits=20, batch=2, seed=13)
    qm1 = ks.prm.p.to(torch.int64) - 1
    ks.x.copy_(qm1.expand_as(ks.x))
    ks.evk.copy_(qm1.expand_as(ks.evk))
    ks.run()
    torch.cuda.synchronize()
    want = ks.out.clone()
    out = torch.zeros_like(want)
    ks.run_fused(out)
    torch.cuda.synchronize()
    assert ks.fused is not None  # the fused kernels took these inputs
    assert torch.equal(out, want)
    if n == 4096:
        assert torch.equal(out[0].cpu(), oracle_keyswitch(ks, 0))
    ks.close()


def test_fused_keyswitch_errors(gpu):
    from heal_amd import capi, params as P

    n, L = 4096, 2
    primes = P.keyswitch_primes(L, n)
    prm = P.NttParams(n, primes, device="cuda")
    fwd = capi.NttPlan(prm.p, prm.perm, prm.twiddles)
    inv = capi.NttPlan(prm.p, prm.perm, prm.inv_twiddles, prm.m_inv, inverse=True)
    gal = P.galois_perms(n, 5, L, device="cuda")
    with pytest.raises(ValueError):  # two forward plans
        capi.KeySwitchPlan(fwd, fwd, gal, 3, 20)
    other = capi.NttPlan(prm.p, torch.arange(n, device="cuda"), prm.inv_twiddles, prm.m_inv, inverse=True)
    with pytest.raises(ValueError):  # different permutations
        capi.KeySwitchPlan(fwd, other, gal, 3, 20)
    bad = gal.clone()
    bad[1, 3] = bad[1, 4]
    with pytest.raises(ValueError):  # a row that is not a permutation
        capi.KeySwitchPlan(fwd, inv, bad, 3, 20)
    bad[1, 3] = n
    with pytest.raises(IndexError):
        capi.KeySwitchPlan(fwd, inv, bad, 3, 20)
    small = P.NttParams(1024, P.keyswitch_primes(L, 1024), device="cuda")
    f2 = capi.NttPlan(small.p, small.perm, small.twiddles)
    i2 = capi.NttPlan(small.p, small.perm, small.inv_twiddles, small.m_inv, inverse=True)
    with pytest.raises(ValueError):  # m below 2^12
        capi.KeySwitchPlan(f2, i2, P.galois_perms(1024, 5, L, device="cuda"), 3, 20)
    for p in (fwd, inv, other, f2, i2):
        p.close()


@pytest.mark.slow
def test_keyswitch_full_size_vs_oracle(gpu):
    """Config: N = 2^16, L = 24 primes of 50-60 bits, dnum = 3; w = 60 (as specified) and w = 20 (all digits live)."""
    from heal_amd.keyswitch import KeySwitchBatch

    for w in (60, 20):
        print(f"[keyswitch-full] w={w}", flush=True)
        ks = KeySwitchBatch(n=65536, limbs=24, base_bits=w, batch=1, seed=w)
        ks.run()
        torch.cuda.synchronize()
        assert torch.equal(ks.out[0].cpu(), oracle_keyswitch(ks, 0))
        ks.close()
