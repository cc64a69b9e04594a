"""The matrix-core NTT (N = 65536, primes < 2^28): bit-exact against the ALU
NTT on the same inputs (limb-major and limb-last layouts, inputs outside
[0, q) included) and against the oracle's radix-2 network on whole columns."""
from __future__ import annotations

import pytest
import torch

from heal_cases import OracleBackend

pytestmark = pytest.mark.gpu

N = 65536


@pytest.fixture(scope="module")
def setup(gpu):
    from heal_amd import capi
    from heal_amd.ntt_mfma import NttMfmaPlan, small_primes
    from heal_amd.params import NttParams

    primes = small_primes(4)
    prm = NttParams(N, primes, device="cuda")
    return primes, prm, NttMfmaPlan(primes, prm.psi), capi.NttPlan(prm.p, prm.perm, prm.twiddles)


def test_matches_alu_ntt_limb_major(setup):
    primes, prm, mfma, alu = setup
    print("[mfma] limb-major", flush=True)
    g = torch.Generator(device="cuda")
    g.manual_seed(1)
    q = prm.p.view(1, -1, 1)
    data = torch.randint(0, 1 << 62, (3, 4, N), generator=g, device="cuda") % q
    data[0, 0, :5] = torch.tensor([-1, -(1 << 40), 1 << 62, primes[0], primes[0] + 3], device="cuda")
    a = data.permute(0, 2, 1).unsqueeze(2)
    o1 = torch.full_like(data, -5).permute(0, 2, 1).unsqueeze(2)
    o2 = torch.full_like(data, -6).permute(0, 2, 1).unsqueeze(2)
    mfma(a, o1)
    alu(a, o2)
    assert torch.equal(o1, o2)


def test_matches_alu_ntt_limb_last_and_in_place(setup):
    primes, prm, mfma, alu = setup
    print("[mfma] limb-last", flush=True)
    g = torch.Generator(device="cuda")
    g.manual_seed(2)
    a = torch.randint(0, 1 << 62, (2, N, 1, 4), generator=g, device="cuda") % prm.p
    exp = torch.empty_like(a)
    alu(a, exp)
    mfma(a, a)  # in place
    assert torch.equal(a, exp)


def test_matches_oracle_columns(setup):
    primes, prm, mfma, alu = setup
    print("[mfma] oracle", flush=True)
    o = OracleBackend()
    g = torch.Generator().manual_seed(3)
    a = torch.randint(0, 1 << 62, (1, N, 1, 4), generator=g) % prm.p.cpu()
    out = torch.empty_like(a).cuda()
    mfma(a.cuda(), out)
    for t in (0, 3):
        exp = torch.zeros(1, N, 1, 1, dtype=torch.int64)
        assert o.ntt(a[..., t:t + 1].contiguous(), prm.p[t:t + 1].cpu(), prm.perm.cpu(),
                     prm.twiddles[t:t + 1].cpu(), exp) is None
        assert torch.equal(out[..., t:t + 1].cpu(), exp)


def test_rejects_large_primes(gpu):
    from heal_amd.ntt_mfma import tables
    from heal_amd.params import keyswitch_primes, primitive_root_2n

    q = keyswitch_primes(1)[0]
    with pytest.raises(ValueError):
        tables([q], [primitive_root_2n(q, N)])


def test_config4_full_size_32_primes(gpu):
    """The exact matrix-core configuration: 16 polynomials x 32 primes below
    2^28, N = 65536, limb-major, seeded uniform residues; bit-exact with the
    ALU NTT on every column and with the oracle's network on two columns."""
    from heal_amd import capi
    from heal_amd.ntt_mfma import NttMfmaPlan, small_primes
    from heal_amd.params import NttParams

    print("[mfma] config 4: 16 x 32 primes", flush=True)
    primes = small_primes(32)
    assert len(primes) == 32 and max(primes) < (1 << 28)
    prm = NttParams(N, primes, device="cuda")
    mfma = NttMfmaPlan(primes, prm.psi)
    alu = capi.NttPlan(prm.p, prm.perm, prm.twiddles)
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    data = torch.randint(0, 1 << 62, (16, 32, N), generator=g, device="cuda") % prm.p.view(1, -1, 1)
    a = data.permute(0, 2, 1).unsqueeze(2)  # [16, N, 1, 32], limb-major in memory
    o1 = torch.zeros_like(data).permute(0, 2, 1).unsqueeze(2)
    o2 = torch.ones_like(data).permute(0, 2, 1).unsqueeze(2)
    mfma(a, o1)
    alu(a, o2)
    torch.cuda.synchronize()
    assert torch.equal(o1, o2)
    o = OracleBackend()
    for b, t in ((0, 0), (15, 31)):
        col = a[b:b + 1, :, :, t:t + 1].cpu().contiguous()
        exp = torch.zeros_like(col)
        assert o.ntt(col, prm.p[t:t + 1].cpu(), prm.perm.cpu(), prm.twiddles[t:t + 1].cpu(), exp) is None
        assert torch.equal(o1[b:b + 1, :, :, t:t + 1].cpu(), exp), (b, t)
    mfma.close()
    alu.close()
