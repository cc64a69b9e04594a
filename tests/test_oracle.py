"""Pins the CPU oracle (oracle/heal_oracle.c) before it is trusted as the checker.

1. every restated reference unit test (tests/golden/reference_kat.json);
2. closed forms: the forward NTT against the defining sum
   out[u] = sum_i a_i psi^((2u+1)i) (tests/golden/ntt_closed_form.json and
   random cases), INTT o NTT = id, modular ops against Python big integers
   with C++ truncation, digits that reconstruct the input, permute against
   numpy take_along_axis;
3. the compiled reference (oracle/_ref, when built in this container) on
   random inputs, including the reference's wrapped INTT scaling at 60-bit
   primes (oracle mode 1).
"""
from __future__ import annotations

import json
import os
import random

import numpy as np
import pytest
import torch

from heal_cases import OracleBackend, RefBackend, check_case, load_cases
from oracle_lib import reference
from heal_amd import params as P

CASES = load_cases()
IDS = [c["name"] for c in CASES]


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_oracle_reproduces_reference_tests(case):
    if check_case(case, OracleBackend()) == "skip":
        pytest.skip("metadata-only view op: covered by the lattica_hw layer")


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_compiled_reference_reproduces_fixtures(case):
    if reference() is None:
        pytest.skip("oracle/_ref not built (needs /root/reference)")
    if check_case(case, RefBackend()) == "skip":
        pytest.skip("metadata-only view op")


# ------------------------------------------------------------ closed forms
def _ntt_oracle(a, prm, dtype=torch.int64):
    o = OracleBackend()
    r = torch.zeros_like(a)
    st = o.ntt(a, prm.p.to(dtype), prm.perm.to(dtype), prm.twiddles.to(dtype), r)
    assert st is None
    return r


def _intt_oracle(a, prm, dtype=torch.int64, mode=0):
    o = OracleBackend()
    r = torch.zeros_like(a)
    st = o.intt(a, prm.p.to(dtype), prm.perm.to(dtype), prm.inv_twiddles.to(dtype), prm.m_inv.to(dtype), r, mode)
    assert st is None
    return r


def test_ntt_closed_form_fixture():
    with open(os.path.join(os.path.dirname(__file__), "golden", "ntt_closed_form.json")) as f:
        cases = json.load(f)
    for c in cases:
        primes = [int(q) for q in c["primes"]]
        prm = P.NttParams(c["m"], primes, psi=[int(g) for g in c["psi"]])
        a = torch.tensor([[int(x) for x in row] for row in c["a"]], dtype=torch.int64).view(1, c["m"], 1, len(primes))
        want = torch.tensor([[int(x) for x in row] for row in c["expected"]], dtype=torch.int64).view(a.shape)
        assert torch.equal(_ntt_oracle(a, prm), want), (c["m"], c["bits"])


def _defining_sum(a_col, q, psi):
    m = len(a_col)
    return [sum(a_col[i] * pow(psi, (2 * u + 1) * i, q) for i in range(m)) % q for u in range(m)]


@pytest.mark.parametrize("m,bits", [(2, 20), (8, 30), (32, 50), (128, 60), (256, 62)])
def test_ntt_defining_sum_random(m, bits):
    rng = random.Random(m * 100 + bits)
    primes = P.ntt_primes(m, [bits, bits - 1])
    prm = P.NttParams(m, primes)
    l, r = 2, 3
    a = torch.tensor([[[[rng.randrange(q) for q in primes] for _ in range(r)] for _ in range(m)] for _ in range(l)],
                     dtype=torch.int64)
    out = _ntt_oracle(a, prm)
    for i in range(l):
        for j in range(r):
            for t, (q, psi) in enumerate(zip(primes, prm.psi)):
                assert out[i, :, j, t].tolist() == _defining_sum(a[i, :, j, t].tolist(), q, psi)


@pytest.mark.parametrize("m,bits,dtype", [(4, 16, torch.int32), (64, 30, torch.int32), (1024, 30, torch.int64),
                                          (4096, 50, torch.int64), (512, 60, torch.int64), (256, 62, torch.int64)])
def test_intt_inverts_ntt(m, bits, dtype):
    primes = P.ntt_primes(m, [bits, bits, bits - 2])
    prm = P.NttParams(m, primes, dtype=dtype)
    g = torch.Generator().manual_seed(m + bits)
    a = torch.stack([torch.randint(0, q, (2, m, 1), generator=g, dtype=torch.int64) for q in primes], dim=-1).to(dtype)
    fwd = _ntt_oracle(a, prm, dtype)
    assert torch.equal(_intt_oracle(fwd, prm, dtype), a)


def test_ntt_unreduced_input_equals_reduced():
    m = 64
    primes = P.ntt_primes(m, [30, 40])
    prm = P.NttParams(m, primes)
    g = torch.Generator().manual_seed(3)
    a = torch.randint(0, 2**62, (1, m, 2, 2), generator=g, dtype=torch.int64)
    red = a % prm.p
    assert torch.equal(_ntt_oracle(a, prm), _ntt_oracle(red, prm))


# ------------------------------------------------------------ modular ops
def _cpp_rem(x: int, p: int) -> int:
    r = abs(x) % abs(p)
    return -r if x < 0 else r


def _wrap(v: int, bits: int) -> int:
    v &= (1 << bits) - 1
    return v - (1 << bits) if v >> (bits - 1) else v


@pytest.mark.parametrize("dtype,bits", [(torch.int32, 32), (torch.int64, 64)])
@pytest.mark.parametrize("op", [0, 1])
def test_modop_signed_bigint(dtype, bits, op):
    rng = random.Random(bits * 10 + op)
    lo, hi = -(1 << (bits - 1)), (1 << (bits - 1)) - 1
    special = [lo, hi, 0, 1, -1, lo + 1]
    shape = (5, 7, 3)
    n = int(np.prod(shape))
    av = [rng.choice(special) if rng.random() < 0.2 else rng.randint(lo, hi) for _ in range(n)]
    bv = [rng.choice(special) if rng.random() < 0.2 else rng.randint(lo, hi) for _ in range(n)]
    pv = [rng.choice([rng.randint(1, hi), -rng.randint(1, hi), rng.randint(2, 1000), lo]) for _ in range(3)]
    a = torch.tensor(av, dtype=dtype).view(shape)
    b = torch.tensor(bv, dtype=dtype).view(shape)
    p = torch.tensor(pv, dtype=dtype)
    r = torch.zeros(shape, dtype=dtype)
    assert OracleBackend().modop(op, a, b, 0, p, 0, r) is None
    for idx in range(n):
        x, y, q = av[idx], bv[idx], pv[idx % 3]
        want = _wrap(_cpp_rem(x * y if op == 0 else x + y, q), bits)
        assert r.view(-1)[idx].item() == want, (x, y, q)


def test_gdecomp_reconstructs_and_counts_overflow():
    rng = random.Random(5)
    w, power = 20, 3
    vals = [rng.randrange(0, 1 << 60) for _ in range(200)] + [(1 << 60) + 5, (1 << 62) + 1]
    a = torch.tensor(vals, dtype=torch.int64)
    r = torch.zeros((len(vals), power), dtype=torch.int64)
    st, ov = OracleBackend().g_decomposition(a, r, power, w)
    assert st is None and ov == 2
    for i, v in enumerate(vals[:200]):
        assert sum(int(r[i, d]) << (w * d) for d in range(power)) == v


def test_permute_matches_take_along_axis():
    g = torch.Generator().manual_seed(9)
    a = torch.randint(0, 1000, (2, 64, 5), generator=g, dtype=torch.int64)
    perms = torch.stack([torch.randperm(64, generator=g) for _ in range(5)]).to(torch.int64)
    r = torch.zeros_like(a)
    assert OracleBackend().permute(a, perms, r, 2, 1) is None
    want = np.take_along_axis(a.numpy(), perms.numpy().T[None, :, :], axis=1)
    assert np.array_equal(r.numpy(), want)


# ------------------------------------------------------------ vs the compiled reference
ref_only = pytest.mark.skipif(reference() is None, reason="oracle/_ref not built (needs /root/reference)")


@ref_only
@pytest.mark.parametrize("dtype", [torch.int32, torch.int64])
def test_oracle_vs_reference_modops_random(dtype):
    g = torch.Generator().manual_seed(11)
    info = torch.iinfo(dtype)
    for trial in range(6):
        shape_r = (3, 4, 5)
        a = torch.randint(info.min, info.max, (3, 1, 5), generator=g, dtype=dtype)
        b = torch.randint(info.min, info.max, (4, 5), generator=g, dtype=dtype)
        p = torch.randint(1, info.max, (5,), generator=g, dtype=dtype) * (1 if trial % 2 else -1)
        for op in (0, 1):
            r1 = torch.zeros(shape_r, dtype=dtype)
            r2 = torch.zeros(shape_r, dtype=dtype)
            assert OracleBackend().modop(op, a, b, 0, p, 0, r1) is None
            assert RefBackend().modop(op, a, b, 0, p, 0, r2) is None
            assert torch.equal(r1, r2)


@ref_only
def test_oracle_vs_reference_axis_gdecomp_permute():
    g = torch.Generator().manual_seed(12)
    for dtype in (torch.int32, torch.int64):
        info = torch.iinfo(dtype)
        p = torch.randint(2, 1 << 20, (6,), generator=g, dtype=dtype)
        a = torch.randint(0, 1 << 20, (4, 3, 6), generator=g, dtype=dtype)
        r1, r2 = torch.zeros((4, 6), dtype=dtype), torch.zeros((4, 6), dtype=dtype)
        assert OracleBackend().axis_modsum(a, p, r1, 1) is None and RefBackend().axis_modsum(a, p, r2, 1) is None
        assert torch.equal(r1, r2)
        x = torch.randint(info.min, info.max, (7, 5), generator=g, dtype=dtype)
        d1, d2 = torch.zeros((7, 5, 4), dtype=dtype), torch.zeros((7, 5, 4), dtype=dtype)
        assert OracleBackend().g_decomposition(x, d1, 4, 7)[0] is None
        assert RefBackend().g_decomposition(x, d2, 4, 7)[0] is None
        assert torch.equal(d1, d2)
        a = torch.randint(0, 100, (3, 8, 4), generator=g, dtype=dtype)
        perms = torch.randint(0, 8, (4, 8), generator=g, dtype=dtype)
        r1, r2 = torch.zeros_like(a), torch.zeros_like(a)
        assert OracleBackend().permute(a, perms, r1, 2, 1) is None and RefBackend().permute(a, perms, r2, 2, 1) is None
        assert torch.equal(r1, r2)


@ref_only
@pytest.mark.parametrize("m,bits", [(16, 30), (256, 30), (64, 60), (1024, 59)])
def test_oracle_vs_reference_ntt_intt(m, bits):
    g = torch.Generator().manual_seed(m + bits)
    primes = P.ntt_primes(m, [bits, bits - 1, bits - 2])
    prm = P.NttParams(m, primes)
    # arbitrary (non-psi-power) tables exercise the network itself, not a closed form
    tw = torch.stack([torch.randint(0, q, (m,), generator=g, dtype=torch.int64) for q in primes])
    perm = torch.randperm(m, generator=g).to(torch.int64)
    a = torch.stack([torch.randint(0, q, (2, m, 3), generator=g, dtype=torch.int64) for q in primes], dim=-1)
    for table in (prm.twiddles, tw):
        r1, r2 = torch.zeros_like(a), torch.zeros_like(a)
        assert OracleBackend().ntt(a, prm.p, perm, table, r1) is None
        assert RefBackend().ntt(a, prm.p, perm, table, r2) is None
        assert torch.equal(r1, r2)
        # inverse: mode 1 restates the reference's T-width scaling, which wraps above ~31.5-bit primes
        mode = 1 if bits > 31 else 0
        i1, i2 = torch.zeros_like(a), torch.zeros_like(a)
        assert OracleBackend().intt(a, prm.p, perm, table, prm.m_inv, i1, mode) is None
        assert RefBackend().intt(a, prm.p, perm, table, prm.m_inv, i2) is None
        assert torch.equal(i1, i2)


# ------------------------------------------------------------ signed NTT inputs
SIGNED = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ntt_signed_kat.json")))


def _signed_case_tensors(c):
    dt = torch.int32 if c["dtype"] == "int32" else torch.int64
    t = lambda v: torch.tensor(v, dtype=dt)  # noqa: E731
    return (t(c["a"]).view(1, 2, 1, 1), t(c["p"]), t(c["perm"]), t(c["tw"]), t(c["itw"]), t(c["m_inv"]),
            t(c["want"]).view(1, 2, 1, 1))


def run_signed_case(backend, c, to=lambda x: x):
    """Runs one hand-worked signed case on a backend; returns the result tensor."""
    a, p, perm, tw, itw, minv, want = _signed_case_tensors(c)
    out = to(torch.zeros_like(a))
    if c["kind"] == "fwd":
        st = backend.ntt(to(a), to(p), to(perm), to(tw), out)
    else:
        st = backend.intt(to(a), to(p), to(perm), to(itw), to(minv), out)
    assert st is None
    return out


@pytest.mark.parametrize("case", SIGNED, ids=[c["name"] for c in SIGNED])
def test_oracle_signed_kat(case):
    """The oracle keeps the reference's negative representatives (hand-worked vectors)."""
    got = run_signed_case(OracleBackend(), case)
    assert got.view(-1).tolist() == case["want"]


@ref_only
@pytest.mark.parametrize("case", SIGNED, ids=[c["name"] for c in SIGNED])
def test_reference_signed_kat(case):
    """The compiled reference gives the hand-worked vectors too (pins the derivation)."""
    got = run_signed_case(RefBackend(), case)
    assert got.view(-1).tolist() == case["want"]


@ref_only
@pytest.mark.parametrize("m,bits,dtype", [(16, 13, torch.int32), (256, 15, torch.int32), (64, 30, torch.int64),
                                          (1024, 30, torch.int64), (256, 59, torch.int64)])
def test_oracle_vs_reference_signed_inputs(m, bits, dtype):
    """Negative inputs (forward) and inputs in [-q, 2q) (inverse), with reduced
    and with signed, unreduced twiddle tables: the oracle equals the compiled
    reference (its wrapped scaling, mode 1, above ~31.5-bit primes)."""
    g = torch.Generator().manual_seed(m * 7 + bits)
    primes = P.ntt_primes(m, [bits, bits - 1])
    prm = P.NttParams(m, primes)
    q = prm.p.view(1, 1, 1, -1)
    big = 2**29 if dtype == torch.int32 else 2**61
    a = (torch.randint(-big, big, (2, m, 2, 2), generator=g, dtype=torch.int64)).to(dtype)
    perm = torch.randperm(m, generator=g).to(dtype)
    tables = [prm.twiddles.to(dtype)]
    signed_tw = prm.twiddles - torch.randint(0, 2, prm.twiddles.shape, generator=g) * prm.p.view(-1, 1)
    tables.append(signed_tw.to(dtype))
    for table in tables:
        r1, r2 = torch.zeros_like(a), torch.zeros_like(a)
        assert OracleBackend().ntt(a, prm.p.to(dtype), perm, table, r1) is None
        assert RefBackend().ntt(a, prm.p.to(dtype), perm, table, r2) is None
        assert torch.equal(r1, r2)
        assert bool((r1 < 0).any())  # the signed representatives are exercised
        x = (torch.randint(-1, 2, a.shape, generator=g) * q + torch.stack(
            [torch.randint(0, int(p), (2, m, 2), generator=g) for p in primes], dim=-1)).to(dtype)
        mode = 1 if bits > 31 else 0
        i1, i2 = torch.zeros_like(x), torch.zeros_like(x)
        assert OracleBackend().intt(x, prm.p.to(dtype), perm, table, prm.m_inv.to(dtype), i1, mode) is None
        assert RefBackend().intt(x, prm.p.to(dtype), perm, table, prm.m_inv.to(dtype), i2) is None
        assert torch.equal(i1, i2)


@pytest.mark.parametrize("m,bits", [(64, 30), (4096, 50)])
def test_oracle_signed_outputs_are_congruent_representatives(m, bits):
    """Signed results are representatives in (-q, q) of the same residues as
    the transform of the reduced input."""
    g = torch.Generator().manual_seed(m + bits)
    primes = P.ntt_primes(m, [bits])
    prm = P.NttParams(m, primes)
    q = int(primes[0])
    a = torch.randint(-(2**40), 2**40, (1, m, 1, 1), generator=g, dtype=torch.int64)
    r_signed, r_red = torch.zeros_like(a), torch.zeros_like(a)
    assert OracleBackend().ntt(a, prm.p, prm.perm, prm.twiddles, r_signed) is None
    assert OracleBackend().ntt(a % q, prm.p, prm.perm, prm.twiddles, r_red) is None
    assert bool((r_signed < 0).any()) and bool((r_signed.abs() < q).all())
    assert torch.equal(r_signed % q, r_red)
