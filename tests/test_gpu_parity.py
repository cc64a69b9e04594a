"""Parity of the MI355X library (through its C ABI) with the oracle.

Bit-exact equality on every case: the restated reference unit tests, random
modular ops with signed extremes and broadcasting, g_decomposition digits and
overflow counts, permute (including aliasing and out-of-range perms), and the
NTT/INTT over sizes 2..2^16, int32/int64, 20..62-bit primes, k-last and
limb-major layouts, identity / bit-reversal / random / non-permutation perms,
arbitrary (non-psi-power) twiddle tables, and reusable plans.
"""
from __future__ import annotations

import pytest
import torch

from heal_cases import GpuBackend, OracleBackend, check_case, load_cases
from heal_amd import params as P

pytestmark = pytest.mark.gpu

CASES = load_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_reference_unit_tests_on_gpu(gpu, case):
    print(f"[case] {case['name']}", flush=True)
    if check_case(case, GpuBackend()) == "skip":
        pytest.skip("metadata-only view op: covered by the lattica_hw layer")


def _both(fn_gpu, fn_oracle, *tensors_cpu):
    dev = [t.cuda() if isinstance(t, torch.Tensor) else t for t in tensors_cpu]
    st_g = fn_gpu(*dev)
    st_o = fn_oracle(*tensors_cpu)
    return st_g, st_o, dev


@pytest.mark.parametrize("dtype", [torch.int32, torch.int64])
@pytest.mark.parametrize("op", [0, 1])
def test_modop_random_signed_broadcast(gpu, dtype, op):
    g = torch.Generator().manual_seed(100 + op)
    info = torch.iinfo(dtype)
    G, O = GpuBackend(), OracleBackend()
    for trial in range(8):
        a = torch.randint(info.min, info.max, (4, 1, 33, 24), generator=g, dtype=dtype)
        b = torch.randint(info.min, info.max, (3, 33, 24), generator=g, dtype=dtype)
        if trial % 2:
            a = a % 1000  # small non-negative values
        p = torch.randint(2, info.max, (24,), generator=g, dtype=dtype)
        if trial == 3:
            p[::3] = -p[::3]
        if trial == 5:
            p[:] = torch.tensor([2, 3, 1024, 65537] * 6, dtype=dtype)
        ro = torch.zeros((4, 3, 33, 24), dtype=dtype)
        rg = ro.cuda()
        assert O.modop(op, a, b, 0, p, 0, ro) is None
        assert G.modop(op, a.cuda(), b.cuda(), 0, p.cuda(), 0, rg) is None
        assert torch.equal(rg.cpu(), ro), trial
        # scalar p / scalar b variants
        ps = int(p[0])
        ro2, rg2 = torch.zeros_like(ro), torch.zeros_like(ro).cuda()
        assert O.modop(op, a, None, 12345, None, ps, ro2) is None
        assert G.modop(op, a.cuda(), None, 12345, None, ps, rg2) is None
        assert torch.equal(rg2.cpu(), ro2)


def test_modop_strided_views_and_aliasing(gpu):
    g = torch.Generator().manual_seed(7)
    q = P.keyswitch_primes(24)
    p = torch.tensor(q, dtype=torch.int64)
    base = torch.stack([torch.randint(0, qq, (64, 3), generator=g) for qq in q], dim=1)  # [64, 24, 3]
    dv = base.permute(2, 0, 1).unsqueeze(1)  # [3, 1, 64, 24] strided digit view
    evk = torch.stack([torch.randint(0, qq, (3, 2, 64), generator=g) for qq in q], dim=-1)
    ro = torch.zeros((3, 2, 64, 24), dtype=torch.int64)
    assert OracleBackend().modop(0, dv, evk, 0, p, 0, ro) is None
    bg = base.cuda()
    rg = torch.zeros_like(ro).cuda()
    assert GpuBackend().modop(0, bg.permute(2, 0, 1).unsqueeze(1), evk.cuda(), 0, p.cuda(), 0, rg) is None
    assert torch.equal(rg.cpu(), ro)
    # in place: result aliases a
    x = evk.clone()
    xg = x.cuda()
    assert OracleBackend().modop(1, x, evk, 0, p, 0, x) is None
    assert GpuBackend().modop(1, xg, evk.cuda(), 0, p.cuda(), 0, xg) is None
    assert torch.equal(xg.cpu(), x)


@pytest.mark.parametrize("dtype", [torch.int32, torch.int64])
def test_mod_remainder_random(gpu, dtype):
    g = torch.Generator().manual_seed(5)
    info = torch.iinfo(dtype)
    a = torch.randint(info.min, info.max, (17, 9), generator=g, dtype=dtype)
    b = torch.randint(info.min, info.max, (17, 9), generator=g, dtype=dtype)
    b[b == 0] = 7
    for variant in ("tt", "tc", "ct"):
        ro = torch.zeros_like(a)
        rg = ro.cuda()
        if variant == "tt":
            so = OracleBackend().mod(a, 0, b, 0, ro, True, True)
            sg = GpuBackend().mod(a.cuda(), 0, b.cuda(), 0, rg, True, True)
        elif variant == "tc":
            so = OracleBackend().mod(a, 0, None, -977, ro, True, False)
            sg = GpuBackend().mod(a.cuda(), 0, None, -977, rg, True, False)
        else:
            so = OracleBackend().mod(None, int(info.min) + 3, b, 0, ro, False, True)
            sg = GpuBackend().mod(None, int(info.min) + 3, b.cuda(), 0, rg, False, True)
        assert so is None and sg is None
        assert torch.equal(rg.cpu(), ro), variant


@pytest.mark.parametrize("dtype", [torch.int32, torch.int64])
def test_axis_modsum_random(gpu, dtype):
    g = torch.Generator().manual_seed(9)
    info = torch.iinfo(dtype)
    for axis in (0, 1, 2):
        a = torch.randint(0, info.max // 8, (3, 5, 7, 24), generator=g, dtype=dtype)
        if axis == 1:
            a = torch.randint(info.min // 8, info.max // 8, (3, 5, 7, 24), generator=g, dtype=dtype)
        p = torch.randint(2, info.max // 4, (24,), generator=g, dtype=dtype)
        shape = list(a.shape)
        del shape[axis]
        ro = torch.zeros(shape, dtype=dtype)
        rg = ro.cuda()
        assert OracleBackend().axis_modsum(a, p, ro, axis) is None
        assert GpuBackend().axis_modsum(a.cuda(), p.cuda(), rg, axis) is None
        assert torch.equal(rg.cpu(), ro), axis


@pytest.mark.parametrize("dtype", [torch.int32, torch.int64])
@pytest.mark.parametrize("w,power", [(1, 6), (20, 3), (60, 3), (0, 2), (63, 2)])
def test_gdecomp_random(gpu, dtype, w, power):
    g = torch.Generator().manual_seed(w * 10 + power)
    info = torch.iinfo(dtype)
    a = torch.randint(info.min, info.max, (33, 24), generator=g, dtype=dtype)
    ro = torch.zeros((33, 24, power), dtype=dtype)
    rg = ro.cuda()
    so, ovo = OracleBackend().g_decomposition(a, ro, power, w)
    sg, ovg = GpuBackend().g_decomposition(a.cuda(), rg, power, w)
    assert so is None and sg is None
    assert torch.equal(rg.cpu(), ro)
    assert ovo == ovg


def test_permute_random_alias_and_oob(gpu):
    g = torch.Generator().manual_seed(13)
    a = torch.randint(0, 2**62, (2, 256, 24), generator=g, dtype=torch.int64)
    perms = P.galois_perms(256, 5, 24)
    ro = torch.zeros_like(a)
    rg = ro.cuda()
    assert OracleBackend().permute(a, perms, ro, 2, 1) is None
    assert GpuBackend().permute(a.cuda(), perms.cuda(), rg, 2, 1) is None
    assert torch.equal(rg.cpu(), ro)
    # result aliases the input
    ag = a.cuda()
    assert GpuBackend().permute(ag, perms.cuda(), ag, 2, 1) is None
    assert torch.equal(ag.cpu(), ro)
    # out-of-range entries raise out_of_range from the call
    bad = perms.clone()
    bad[3, 17] = 256
    assert GpuBackend().permute(a.cuda(), bad.cuda(), torch.zeros_like(a).cuda(), 2, 1) == "out_of_range"
    bad[3, 17] = -1
    assert GpuBackend().permute(a.cuda(), bad.cuda(), torch.zeros_like(a).cuda(), 2, 1) == "out_of_range"


def test_permute_rows_same_flag(gpu):
    """A table whose rows all equal row 0 (one Galois map for every limb) is
    reported as such, and the gather that reads row 0 only equals the oracle;
    tables with one differing entry or per-row permutations are not."""
    from heal_amd import capi

    g = torch.Generator().manual_seed(17)
    a = torch.randint(-2**62, 2**62, (4, 512, 24), generator=g, dtype=torch.int64)
    perms = P.galois_perms(512, 5, 24)
    want = torch.zeros_like(a)
    assert OracleBackend().permute(a, perms, want, 2, 1) is None
    pg = perms.cuda()
    assert capi.perms_rows_same(pg)
    out = torch.zeros_like(a).cuda()
    capi.permute(a.cuda(), pg, out, 2, 1, trusted=True, rows_same=True)
    assert torch.equal(out.cpu(), want)
    one_off = perms.clone()
    one_off[23, 511], one_off[23, 510] = perms[23, 510], perms[23, 511]
    assert not capi.perms_rows_same(one_off.cuda())
    rows = torch.stack([torch.randperm(512, generator=g) for _ in range(24)]).to(torch.int64)
    assert not capi.perms_rows_same(rows.cuda())
    # a strided (transposed-back) view of the same table and a one-row table
    assert capi.perms_rows_same(pg.t().contiguous().t())
    assert capi.perms_rows_same(pg[:1])
    # int32 tensors and table
    a32 = torch.randint(-2**31, 2**31, (3, 512, 24), generator=g, dtype=torch.int32)
    want32 = torch.zeros_like(a32)
    assert OracleBackend().permute(a32, perms.to(torch.int32), want32, 2, 1) is None
    pg32 = pg.to(torch.int32)
    assert capi.perms_rows_same(pg32)
    assert not capi.perms_rows_same(one_off.to(torch.int32).cuda())
    out32 = torch.zeros_like(a32).cuda()
    capi.permute(a32.cuda(), pg32, out32, 2, 1, trusted=True, rows_same=True)
    assert torch.equal(out32.cpu(), want32)


# ------------------------------------------------------------------- NTT
def _ntt_case(m, primes, dtype, layout, perm_kind, table_kind, gen, l=2, r=3):
    prm = P.NttParams(m, primes, dtype=dtype)
    k = len(primes)
    if perm_kind == "bitrev":
        perm = prm.perm
    elif perm_kind == "identity":
        perm = torch.arange(m, dtype=dtype)
    elif perm_kind == "random":
        perm = torch.randperm(m, generator=gen).to(dtype)
    else:  # not a permutation
        perm = torch.randint(0, m, (m,), generator=gen).to(dtype)
    if table_kind == "psi":
        tw, itw = prm.twiddles, prm.inv_twiddles
    else:
        tw = torch.stack([torch.randint(0, q, (m,), generator=gen) for q in primes]).to(dtype)
        itw = torch.stack([torch.randint(0, q, (m,), generator=gen) for q in primes]).to(dtype)
    data = torch.stack([torch.randint(0, q, (l, m, r), generator=gen) for q in primes], dim=-1).to(dtype)
    if layout == "klast":
        a = data
    else:  # limb-major storage [k, l, r, m] viewed as [l, m, r, k]
        store = data.permute(3, 0, 2, 1).contiguous()
        a = store.permute(1, 3, 2, 0)
    return prm, perm, tw, itw, a


NTT_GRID = [
    # m, bits, dtype, layout, perm, table
    (2, [20], torch.int32, "klast", "bitrev", "psi"),
    (4, [17], torch.int32, "klast", "bitrev", "psi"),
    (16, [30, 28], torch.int32, "klast", "bitrev", "psi"),
    (64, [60, 50], torch.int64, "klast", "bitrev", "psi"),
    (256, [62, 31, 30], torch.int64, "klast", "random", "psi"),
    (1024, [45, 60], torch.int64, "limbmajor", "bitrev", "psi"),
    (4096, [30, 30, 30, 30], torch.int64, "limbmajor", "bitrev", "psi"),
    (4096, [59, 58], torch.int64, "klast", "identity", "random"),
    (8192, [60, 36], torch.int64, "klast", "bitrev", "psi"),
    (16384, [30, 55], torch.int64, "limbmajor", "random", "psi"),
    (32768, [61, 29], torch.int64, "klast", "bitrev", "random"),
    (65536, [60, 50], torch.int64, "klast", "bitrev", "psi"),
    (65536, [30, 60], torch.int64, "limbmajor", "bitrev", "psi"),
    (128, [40], torch.int64, "klast", "nonperm", "psi"),
    (8192, [40, 20], torch.int64, "klast", "nonperm", "random"),
]


@pytest.mark.parametrize("m,bits,dtype,layout,perm_kind,table", NTT_GRID,
                         ids=[f"m{c[0]}-{c[3]}-{c[4]}-{c[5]}-{str(c[2])[6:]}" for c in NTT_GRID])
def test_ntt_intt_parity(gpu, m, bits, dtype, layout, perm_kind, table):
    gen = torch.Generator().manual_seed(m + len(bits))
    primes = P.ntt_primes(m, bits)
    prm, perm, tw, itw, a = _ntt_case(m, primes, dtype, layout, perm_kind, table, gen)
    print(f"[ntt] m={m} bits={bits} {layout} {perm_kind} {table}", flush=True)
    O, G = OracleBackend(), GpuBackend()
    # forward
    ro = torch.zeros_like(a)
    assert O.ntt(a, prm.p, perm, tw, ro) is None
    ag = a.cuda() if layout == "klast" else a.permute(3, 0, 2, 1).contiguous().cuda().permute(1, 3, 2, 0)
    rg = torch.zeros(a.shape, dtype=dtype, device="cuda")
    assert G.ntt(ag, prm.p.cuda(), perm.cuda(), tw.cuda(), rg) is None
    assert torch.equal(rg.cpu(), ro)
    # inverse (result starts with non-zero contents: non-permutation perms keep untouched rows)
    init = torch.randint(0, 1000, a.shape, dtype=dtype)
    io = init.clone()
    ig = init.cuda()
    assert O.intt(ro, prm.p, perm, itw, prm.m_inv, io) is None
    assert G.intt(rg, prm.p.cuda(), perm.cuda(), itw.cuda(), prm.m_inv.cuda(), ig) is None
    assert torch.equal(ig.cpu(), io)
    if perm_kind in ("bitrev", "identity", "random") and table == "psi":
        assert torch.equal(ig.cpu(), a)  # round trip


EXTREME_CLASSES = [
    # bits, dtype: the lazy 60-bit class at its largest moduli, the exact 64-bit class, 32-bit Shoup
    ([60, 60], torch.int64),
    ([62, 61], torch.int64),
    ([30, 29], torch.int32),
]


@pytest.mark.parametrize("m", [4096, 65536])
@pytest.mark.parametrize("pattern", ["max", "alternate"])
@pytest.mark.parametrize("cls", range(len(EXTREME_CLASSES)), ids=["lazy60", "exact64", "s32"])
def test_ntt_intt_extreme_residues_at_the_lazy_bound(gpu, m, pattern, cls):
    """Every input q - 1 (or q - 1 and 0 alternating), with the psi tables and
    with every twiddle q - 1, in each butterfly class: at the two largest primes
    below 2^60 the lazy 60-bit butterflies' intermediates reach their bounds
    (below 16q forward, 4q inverse); 61/62-bit primes take the exact 64-bit
    butterflies, 29/30-bit int32 the 32-bit ones.  Forward and inverse equal the
    oracle's."""
    bits, dtype = EXTREME_CLASSES[cls]
    primes = P.ntt_primes(m, bits)
    if cls == 0:
        assert all((1 << 59) < q < (1 << 60) for q in primes)
    O, G = OracleBackend(), GpuBackend()
    prm = P.NttParams(m, primes, dtype=dtype)
    qv = torch.tensor(primes, dtype=dtype)
    a = (qv - 1).view(1, 1, 1, -1).expand(2, m, 1, len(primes)).contiguous()
    if pattern == "alternate":
        a[:, 1::2] = 0
    for table in ("psi", "max"):
        print(f"[ntt extreme] m={m} {bits} {pattern} {table}", flush=True)
        if table == "psi":
            tw, itw = prm.twiddles, prm.inv_twiddles
        else:
            tw = (qv - 1).view(-1, 1).expand(len(primes), m).contiguous()
            itw = tw.clone()
        ro = torch.zeros_like(a)
        assert O.ntt(a, prm.p, prm.perm, tw, ro) is None
        rg = torch.zeros(a.shape, dtype=dtype, device="cuda")
        assert G.ntt(a.cuda(), prm.p.cuda(), prm.perm.cuda(), tw.cuda(), rg) is None
        assert torch.equal(rg.cpu(), ro), table
        io = torch.zeros_like(a)
        ig = torch.zeros(a.shape, dtype=dtype, device="cuda")
        assert O.intt(ro, prm.p, prm.perm, itw, prm.m_inv, io) is None
        assert G.intt(rg, prm.p.cuda(), prm.perm.cuda(), itw.cuda(), prm.m_inv.cuda(), ig) is None
        assert torch.equal(ig.cpu(), io), table
        if table == "psi":
            assert torch.equal(io, a)  # round trip


LIMB_MAJOR_GRID = [
    # m, bits, dtype, perm (input and result both limb-major)
    (4096, [30, 29, 28], torch.int32, "bitrev"),
    (4096, [60, 30], torch.int64, "bitrev"),
    (32768, [59, 27], torch.int64, "bitrev"),
    (65536, [60, 28, 50], torch.int64, "bitrev"),
    (65536, [60, 45], torch.int64, "random"),
]


@pytest.mark.parametrize("m,bits,dtype,perm_kind", LIMB_MAJOR_GRID,
                         ids=[f"m{c[0]}-{c[3]}-{str(c[2])[6:]}" for c in LIMB_MAJOR_GRID])
def test_ntt_intt_limb_major_results(gpu, m, bits, dtype, perm_kind):
    """Limb-major input and result views: the pass at the permuted rows takes
    tiles of bit-reversed rows (bit-reversal perm) or scatters (other perms).
    The int64 inverse also runs on unreduced entries (x + q), where the
    reference's truncating remainders can leave negative representatives: the
    result equals the oracle's, which keeps them."""
    gen = torch.Generator().manual_seed(m + 7)
    primes = P.ntt_primes(m, bits)
    prm, perm, tw, itw, a = _ntt_case(m, primes, dtype, "limbmajor", perm_kind, "psi", gen, l=3, r=1)
    print(f"[ntt limb-major] m={m} bits={bits} {perm_kind} {dtype}", flush=True)
    O, G = OracleBackend(), GpuBackend()

    def limb_major_cuda(t):  # [l, m, r, k] view over [k, l, r, m] storage
        return t.permute(3, 0, 2, 1).contiguous().cuda().permute(1, 3, 2, 0)

    ro = torch.zeros_like(a)
    assert O.ntt(a, prm.p, perm, tw, ro) is None
    rg = limb_major_cuda(torch.zeros_like(a))
    assert G.ntt(limb_major_cuda(a), prm.p.cuda(), perm.cuda(), tw.cuda(), rg) is None
    assert torch.equal(rg.cpu(), ro)
    # inverse from limb-major input, against the oracle and as a round trip
    io = torch.zeros_like(a)
    assert O.intt(ro, prm.p, perm, itw, prm.m_inv, io) is None
    ig = limb_major_cuda(torch.zeros_like(a))
    assert G.intt(limb_major_cuda(ro), prm.p.cuda(), perm.cuda(), itw.cuda(), prm.m_inv.cuda(), ig) is None
    assert torch.equal(ig.cpu(), io)
    # entries in [q, 2q): the oracle's signed network (int64; the int32 sum
    # q + x could pass 2^31 at these primes)
    if dtype == torch.int64:
        x = ro.clone()
        x[:, ::7] += prm.p.view(1, 1, 1, -1).expand_as(x[:, ::7])
        xo = torch.zeros_like(a)
        assert O.intt(x, prm.p, perm, itw, prm.m_inv, xo) is None
        iu = limb_major_cuda(torch.zeros_like(a))
        assert G.intt(limb_major_cuda(x), prm.p.cuda(), perm.cuda(), itw.cuda(), prm.m_inv.cuda(), iu) is None
        assert torch.equal(iu.cpu(), xo)
        assert torch.equal(xo % prm.p, io)  # congruent to the transform of the residues
    assert torch.equal(ig.cpu(), a)  # round trip


def test_ntt_plan_reuse_and_inplace(gpu):
    from heal_amd import capi

    m = 65536
    primes = P.keyswitch_primes(4)
    prm = P.NttParams(m, primes).to("cuda")
    g = torch.Generator().manual_seed(1)
    fwd = capi.NttPlan(prm.p, prm.perm, prm.twiddles)
    inv = capi.NttPlan(prm.p, prm.perm, prm.inv_twiddles, prm.m_inv, inverse=True)
    for trial in range(3):
        a = torch.stack([torch.randint(0, q, (1, m, 2), generator=g) for q in primes], dim=-1)
        x = a.cuda()
        fwd(x, x)  # in place
        ro = torch.zeros_like(a)
        assert OracleBackend().ntt(a, prm.p.cpu(), prm.perm.cpu(), prm.twiddles.cpu(), ro) is None
        assert torch.equal(x.cpu(), ro)
        inv(x, x)
        assert torch.equal(x.cpu(), a)
    fwd.close()
    inv.close()


def test_ntt_unreduced_and_errors(gpu):
    m = 256
    primes = P.ntt_primes(m, [40, 41])
    prm = P.NttParams(m, primes).to("cuda")
    g = torch.Generator().manual_seed(2)
    a = torch.randint(0, 2**62, (1, m, 1, 2), generator=g, dtype=torch.int64)
    ro = torch.zeros_like(a)
    assert OracleBackend().ntt(a, prm.p.cpu(), prm.perm.cpu(), prm.twiddles.cpu(), ro) is None
    rg = torch.zeros_like(a).cuda()
    G = GpuBackend()
    assert G.ntt(a.cuda(), prm.p, prm.perm, prm.twiddles, rg) is None
    assert torch.equal(rg.cpu(), ro)
    bad = prm.perm.clone()
    bad[5] = m
    assert G.ntt(a.cuda(), prm.p, bad, prm.twiddles, rg) == "out_of_range"
    assert G.ntt(a.cuda(), prm.p, prm.perm, prm.twiddles[:, :128].contiguous(), rg) == "invalid_argument"
    assert G.ntt(a.cuda().view(m, 2), prm.p, prm.perm, prm.twiddles, rg.view(m, 2)) == "invalid_argument"


@pytest.mark.slow
def test_ntt_full_keyswitch_shape(gpu):
    """All 72 digit limb-NTTs of one N=2^16, L=24 key-switch (strided digit view) and the 48 INTTs."""
    from heal_amd import capi

    N, L, D = 65536, 24, 3
    primes = P.keyswitch_primes(L)
    prm = P.NttParams(N, primes)
    g = torch.Generator().manual_seed(21)
    digits = torch.stack([torch.randint(0, 2**20, (N, D), generator=g) for _ in range(L)], dim=1)  # [N, L, D]
    dv = digits.view(1, N, L, D).permute(0, 1, 3, 2)  # [1, N, D, L]
    ro = torch.zeros((1, N, D, L), dtype=torch.int64)
    assert OracleBackend().ntt(dv, prm.p, prm.perm, prm.twiddles, ro) is None
    dg = digits.cuda().view(1, N, L, D).permute(0, 1, 3, 2)
    rg = torch.zeros((1, N, D, L), dtype=torch.int64, device="cuda")
    capi.ntt(dg, prm.p.cuda(), prm.perm.cuda(), prm.twiddles.cuda(), rg)
    assert torch.equal(rg.cpu(), ro)
    acc = torch.stack([torch.randint(0, q, (2, N, 1), generator=g) for q in primes], dim=-1)
    io = torch.zeros_like(acc)
    assert OracleBackend().intt(acc, prm.p, prm.perm, prm.inv_twiddles, prm.m_inv, io) is None
    ig = torch.zeros_like(acc).cuda()
    capi.intt(acc.cuda(), prm.p.cuda(), prm.perm.cuda(), prm.inv_twiddles.cuda(), prm.m_inv.cuda(), ig)
    assert torch.equal(ig.cpu(), io)
