"""The C++ HEAL API (lattica_hw_api templates) through its pybind11 module
`lattica_hw`, on the GPU: every restated reference unit test, plus the
behaviour of the C++ layer itself (views sharing one buffer, strides kept
across host/device transfers, the plan cache and its invalidation by writes,
permute validation, the deferred g_decomposition warning)."""
from __future__ import annotations

import pytest
import torch

from heal_cases import LatticaBackend, OracleBackend, check_case, load_cases
from heal_amd import params as P

pytestmark = pytest.mark.gpu

CASES = load_cases()


@pytest.fixture(scope="module")
def lhw(gpu):
    from heal_amd import lattica_hw

    return lattica_hw


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_reference_unit_tests_through_lattica_hw(lhw, case):
    print(f"[lattica_hw] {case['name']}", flush=True)
    assert check_case(case, LatticaBackend()) == "ok"


def test_module_exports_reference_names(lhw):
    names = ["DeviceTensor32", "DeviceTensor64", "DeviceTensorfloat64"]
    for s in ("32", "64", "float64"):
        names += [f"{n}_{s}" for n in ("allocate_on_hardware", "host_to_device", "device_to_host", "expand", "squeeze",
                                      "unsqueeze", "make_contiguous")]
    for s in ("32", "64"):
        names += [f"{n}_{s}" for n in ("modmul_ttt", "modmul_ttc", "modmul_tct", "modmul_tcc", "modsum_ttt",
                                      "modsum_ttc", "modsum_tct", "modsum_tcc", "axis_modsum", "g_decomposition",
                                      "ntt", "intt", "permute", "mod_tt", "mod_tc", "mod_ct")]
    missing = [n for n in names if not hasattr(lhw, n)]
    assert not missing, missing


def test_strides_survive_transfers_and_views(lhw):
    base = torch.arange(24, dtype=torch.int64).reshape(2, 3, 4)
    t = base.transpose(0, 2)  # non-contiguous
    hw = lhw.host_to_device_64(t)
    assert hw.strides == list(t.stride()) and not hw.is_contiguous()
    assert torch.equal(lhw.device_to_host_64(hw), t)
    same = lhw.make_contiguous_64(hw)
    assert same is hw and hw.is_contiguous()
    assert torch.equal(lhw.device_to_host_64(hw), t)
    # expand + broadcast in modmul
    col = lhw.host_to_device_64(torch.tensor([[3], [5]], dtype=torch.int64))
    lhw.expand_64(col, 1, 4)
    assert col.shape == [2, 4] and col.strides[1] == 0
    out = lhw.allocate_on_hardware_64([2, 4])
    b = lhw.host_to_device_64(torch.arange(8, dtype=torch.int64).reshape(2, 4))
    p = lhw.host_to_device_64(torch.tensor([7, 11, 13, 17], dtype=torch.int64))
    lhw.modmul_ttt_64(col, b, p, out)
    want = (torch.tensor([[3], [5]]) * torch.arange(8).reshape(2, 4)) % torch.tensor([7, 11, 13, 17])
    assert torch.equal(lhw.device_to_host_64(out), want)


def test_ntt_plan_cache_and_invalidation(lhw):
    m = 4096
    primes = P.ntt_primes(m, [50, 40])
    prm = P.NttParams(m, primes)
    g = torch.Generator().manual_seed(4)
    a = torch.stack([torch.randint(0, q, (2, m, 1), generator=g) for q in primes], dim=-1)
    hw = {k: lhw.host_to_device_64(getattr(prm, k)) for k in ("p", "perm", "twiddles", "inv_twiddles", "m_inv")}
    A = lhw.host_to_device_64(a)
    R = lhw.allocate_on_hardware_64(list(a.shape))
    before = lhw.plan_cache_size()
    lhw.ntt_64(A, hw["p"], hw["perm"], hw["twiddles"], None, None, R)
    mid = lhw.plan_cache_size()
    lhw.ntt_64(A, hw["p"], hw["perm"], hw["twiddles"], None, None, R)
    assert lhw.plan_cache_size() == mid == before + 1  # the second call reuses the plan
    want = torch.zeros_like(a)
    assert OracleBackend().ntt(a, prm.p, prm.perm, prm.twiddles, want) is None
    assert torch.equal(lhw.device_to_host_64(R), want)
    # writing into the twiddle tensor (any op's result) invalidates the plan
    tw2 = prm.twiddles.clone()
    tw2[:, 5] = 12345
    src = lhw.host_to_device_64(tw2)
    ones = lhw.host_to_device_64(torch.ones(1, dtype=torch.int64))
    big = lhw.host_to_device_64(torch.tensor([2 ** 62], dtype=torch.int64))
    lhw.modmul_ttt_64(src, ones, big, hw["twiddles"])  # twiddles := tw2
    lhw.ntt_64(A, hw["p"], hw["perm"], hw["twiddles"], None, None, R)
    want2 = torch.zeros_like(a)
    assert OracleBackend().ntt(a, prm.p, prm.perm, tw2, want2) is None
    assert torch.equal(lhw.device_to_host_64(R), want2)
    # round trip through the inverse
    lhw.intt_64(lhw.host_to_device_64(want), hw["p"], hw["perm"], hw["inv_twiddles"], hw["m_inv"], None, None, R)
    assert torch.equal(lhw.device_to_host_64(R), a)


def test_permute_validation_cache(lhw):
    a = torch.randint(0, 100, (2, 8, 3), dtype=torch.int64)
    perms = torch.stack([torch.randperm(8) for _ in range(3)]).to(torch.int64)
    A, PM = lhw.host_to_device_64(a), lhw.host_to_device_64(perms)
    R = lhw.allocate_on_hardware_64([2, 8, 3])
    for _ in range(2):
        lhw.permute_64(A, PM, R, 2, 1)
        want = torch.zeros_like(a)
        assert OracleBackend().permute(a, perms, want, 2, 1) is None
        assert torch.equal(lhw.device_to_host_64(R), want)
    bad = perms.clone()
    bad[0, 0] = 8
    with pytest.raises(IndexError):
        lhw.permute_64(A, lhw.host_to_device_64(bad), R, 2, 1)


def test_permute_cache_identical_rows(lhw):
    """The validation cache records whether every row equals row 0; the later
    (trusted) calls then read row 0 only and still equal the oracle.  A table
    with per-row permutations in the same session keeps reading every row."""
    g = torch.Generator().manual_seed(5)
    a = torch.randint(0, 2**60, (2, 64, 6), generator=g, dtype=torch.int64)
    same = P.galois_perms(64, 5, 6)
    rows = torch.stack([torch.randperm(64, generator=g) for _ in range(6)]).to(torch.int64)
    A = lhw.host_to_device_64(a)
    R = lhw.allocate_on_hardware_64([2, 64, 6])
    for perms in (same, rows):
        PM = lhw.host_to_device_64(perms)
        want = torch.zeros_like(a)
        assert OracleBackend().permute(a, perms, want, 2, 1) is None
        for _ in range(3):
            lhw.permute_64(A, PM, R, 2, 1)
            assert torch.equal(lhw.device_to_host_64(R), want)
    # the identical-rows table rewritten in place with per-row maps: a new
    # version, validated again, read row by row
    PM = lhw.host_to_device_64(same)
    lhw.permute_64(A, PM, R, 2, 1)
    lhw.permute_64(A, PM, R, 2, 1)
    lhw.copy_from_host_64(PM, rows)
    want = torch.zeros_like(a)
    assert OracleBackend().permute(a, rows, want, 2, 1) is None
    for _ in range(2):
        lhw.permute_64(A, PM, R, 2, 1)
        assert torch.equal(lhw.device_to_host_64(R), want)


def test_gdecomp_overflow_warning_reported_on_sync(lhw, capfd):
    """The reference's per-element warning (example_impl/g_decomposition_impl.cpp:58-65),
    printed at the next synchronisation."""
    a = lhw.host_to_device_64(torch.tensor([5, 1 << 40, 7], dtype=torch.int64))
    r = lhw.allocate_on_hardware_64([3, 2])
    lhw.g_decomposition_64(a, r, 2, 8)
    out = lhw.device_to_host_64(r)
    assert out.tolist() == [[5, 0], [0, 0], [7, 0]]
    err = capfd.readouterr().err
    assert err == "Warning: value at 1 exceeds capacity with base_bits=8 and power=2\n"
    # two calls of other shapes before one synchronisation: each element with its call's coordinates
    x = torch.zeros(2, 3, dtype=torch.int64)
    x[1, 2] = 1 << 20
    x[0, 1] = -1  # size_t digit arithmetic: -1 leaves 2^56 - 1 after two 4-bit digits
    b = lhw.host_to_device_64(x)
    rb = lhw.allocate_on_hardware_64([2, 3, 2])
    lhw.g_decomposition_64(b, rb, 2, 4)
    c = lhw.host_to_device_64(torch.tensor([[1 << 30, 1], [3, 1 << 30]], dtype=torch.int64))
    rc = lhw.allocate_on_hardware_64([2, 2, 3])
    lhw.g_decomposition_64(c, rc, 3, 5)
    lhw.device_to_host_64(rc)
    err = capfd.readouterr().err.splitlines()
    assert err == ["Warning: value at 0 1 exceeds capacity with base_bits=4 and power=2",
                   "Warning: value at 1 2 exceeds capacity with base_bits=4 and power=2",
                   "Warning: value at 0 0 exceeds capacity with base_bits=5 and power=3",
                   "Warning: value at 1 1 exceeds capacity with base_bits=5 and power=3"], err
    lhw.device_to_host_64(rc)
    assert capfd.readouterr().err == ""


def test_allocate_on_hardware_returns_zeros(lhw):
    """allocate_on_hardware zero-fills its buffer (example_impl/device_memory_impl.cpp:168-179),
    also when the stream-ordered pool hands back memory that held other values."""
    for suffix, dtype in (("32", torch.int32), ("64", torch.int64), ("float64", torch.float64)):
        for shape in ([1 << 20], [3, 5, 7], [2, 0, 4], []):
            n = 1
            for d in shape:
                n *= d
            junk = getattr(lhw, f"host_to_device_{suffix}")(torch.full((max(n, 1) + 13,), 7, dtype=dtype))
            del junk  # its memory returns to the pool
            t = getattr(lhw, f"allocate_on_hardware_{suffix}")(shape)
            host = getattr(lhw, f"device_to_host_{suffix}")(t)
            assert list(host.shape) == shape
            assert bool((host == 0).all()), (suffix, shape)
