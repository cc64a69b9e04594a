"""Transcript operations that the reference runtime names without defining
(python_execution/lattica_heal_runtime/device_interface.py: _modneg_*,
modmul_axis_sum, abs, set_const_val, moveaxis, get_slice, pad_single_axis), on
the GPU through the C ABI, the lattica_hw module and the transcript dispatcher.

modmul_axis_sum fuses two reference ops, so it is checked against the oracle's
modmul_ttt followed by its axis_modsum (the chain it replaces).  The other ops
have no reference definition; they are checked against their closed forms
(Python integers, torch indexing on the CPU)."""
from __future__ import annotations

import pytest
import torch

from heal_cases import OracleBackend

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def capi(gpu):
    from heal_amd import capi

    return capi


@pytest.fixture(scope="module")
def lhw(gpu):
    from heal_amd import lattica_hw

    return lattica_hw


@pytest.fixture(scope="module")
def engine(gpu):
    from heal_amd.runtime.dispatcher import HealDispatcher

    return HealDispatcher()


def _rand(shape, dtype, g, lo=None, hi=None):
    info = torch.iinfo(dtype)
    lo = info.min if lo is None else lo
    hi = info.max if hi is None else hi
    return torch.randint(lo, hi, shape, generator=g, dtype=torch.int64).to(dtype)


def _neg_residue(a, p):
    P = abs(int(p))
    return 0 if P == 0 else (-int(a)) % P


@pytest.mark.parametrize("dtype", [torch.int32, torch.int64])
def test_modneg_closed_form(capi, dtype):
    print(f"[ext] modneg {dtype}", flush=True)
    g = torch.Generator().manual_seed(11)
    a = _rand((4, 33, 7), dtype, g)
    a[0, 0, 0] = torch.iinfo(dtype).min
    a[0, 0, 1] = 0
    p = _rand((7,), dtype, g, -(1 << 29), 1 << 29)
    p[3] = 0
    p[4] = 1
    p[5] = -(1 << 20) - 7
    exp = torch.tensor([[[_neg_residue(a[i, j, k], p[k]) for k in range(7)] for j in range(33)] for i in range(4)],
                       dtype=dtype)
    out = torch.full_like(a, 99).cuda()
    capi.modneg_tt(a.cuda(), p.cuda(), out)
    assert torch.equal(out.cpu(), exp)
    exp_c = torch.tensor([[[_neg_residue(x, 1000003) for x in row] for row in m] for m in a.tolist()], dtype=dtype)
    capi.modneg_tc(a.cuda(), 1000003, out)
    assert torch.equal(out.cpu(), exp_c)


def _oracle_chain(a, b, p, axis):
    """modmul_ttt into the broadcast shape, then axis_modsum: the oracle's chain."""
    o = OracleBackend()
    shape = torch.broadcast_shapes(a.shape, b.shape)
    prod = torch.zeros(shape, dtype=a.dtype)
    assert o.modop(0, a, b, 0, p, 0, prod) is None  # None: success
    rshape = list(shape)
    del rshape[axis]
    out = torch.zeros(rshape, dtype=a.dtype)
    assert o.axis_modsum(prod, p, out, axis) is None
    return out


@pytest.mark.parametrize("dtype", [torch.int32, torch.int64])
@pytest.mark.parametrize("signed", [False, True])
def test_modmul_axis_sum_matches_oracle_chain(capi, lhw, engine, dtype, signed):
    print(f"[ext] modmul_axis_sum {dtype} signed={signed}", flush=True)
    g = torch.Generator().manual_seed(12 + signed)
    bits = 30 if dtype == torch.int32 else 60
    p = torch.randint(1 << (bits - 2), 1 << (bits - 1), (5,), generator=g, dtype=torch.int64).to(dtype)
    if signed:
        a = _rand((3, 1, 64, 5), dtype, g)
        b = _rand((3, 2, 64, 5), dtype, g)
        p[1] = -p[1]
    else:  # the key-switch case: residues times key words in [0, q)
        a = (torch.randint(0, 1 << 62, (3, 1, 64, 5), generator=g) % p.to(torch.int64)).to(dtype)
        b = (torch.randint(0, 1 << 62, (3, 2, 64, 5), generator=g) % p.to(torch.int64)).to(dtype)
    exp = _oracle_chain(a, b, p, 0)
    out = torch.zeros(2, 64, 5, dtype=dtype).cuda()
    capi.modmul_axis_sum(a.cuda(), b.cuda(), p.cuda(), out, 0)
    assert torch.equal(out.cpu(), exp)
    # axis 1 of the broadcast shape, b broadcast instead of a
    exp1 = _oracle_chain(b, a, p, 1)
    out1 = torch.zeros(3, 64, 5, dtype=dtype).cuda()
    capi.modmul_axis_sum(b.cuda(), a.cuda(), p.cuda(), out1, 1)
    assert torch.equal(out1.cpu(), exp1)
    # the C++ template through lattica_hw, and the transcript op
    sfx = "32" if dtype == torch.int32 else "64"
    up = getattr(lhw, f"host_to_device_{sfx}")
    res = getattr(lhw, f"allocate_on_hardware_{sfx}")([2, 64, 5])
    getattr(lhw, f"modmul_axis_sum_{sfx}")(up(a), up(b), up(p), res, 0)
    assert torch.equal(getattr(lhw, f"device_to_host_{sfx}")(res), exp)
    res2 = getattr(lhw, f"allocate_on_hardware_{sfx}")([2, 64, 5])
    engine.modmul_axis_sum(up(a), up(b), up(p), res2, 0)
    assert torch.equal(engine.device_to_host(res2), exp)


def test_modmul_axis_sum_errors(capi):
    a = torch.zeros(3, 4, 5, dtype=torch.int64).cuda()
    p = torch.ones(5, dtype=torch.int64).cuda()
    with pytest.raises(ValueError):
        capi.modmul_axis_sum(a, a, p, torch.zeros(3, 4, dtype=torch.int64).cuda(), 2)  # the limb axis
    with pytest.raises(ValueError):
        capi.modmul_axis_sum(a, torch.zeros(2, 4, 5, dtype=torch.int64).cuda(), p,
                             torch.zeros(4, 5, dtype=torch.int64).cuda(), 0)
    with pytest.raises(ValueError):
        capi.modmul_axis_sum(a, a, torch.ones(4, dtype=torch.int64).cuda(), torch.zeros(4, 5, dtype=torch.int64).cuda(), 0)


@pytest.mark.parametrize("dtype", [torch.int32, torch.int64])
def test_abs_and_set_const(capi, lhw, dtype):
    print(f"[ext] abs/set_const {dtype}", flush=True)
    g = torch.Generator().manual_seed(13)
    a = _rand((6, 40), dtype, g)
    a[0, 0] = torch.iinfo(dtype).min
    out = torch.zeros_like(a).cuda()
    capi.abs_(a.cuda(), out)
    assert torch.equal(out.cpu(), torch.abs(a))  # the minimum stays the minimum
    view = out[:, 3:17:2]
    capi.set_const(view, -5)
    exp = torch.abs(a)
    exp[:, 3:17:2] = -5
    assert torch.equal(out.cpu(), exp)
    d = torch.zeros(4, 3, dtype=torch.float64).cuda()
    capi.set_const(d, 2.5)
    assert torch.equal(d.cpu(), torch.full((4, 3), 2.5, dtype=torch.float64))
    sfx = "32" if dtype == torch.int32 else "64"
    t = getattr(lhw, f"host_to_device_{sfx}")(a)
    r = getattr(lhw, f"allocate_on_hardware_{sfx}")([6, 40])
    getattr(lhw, f"abs_{sfx}")(t, r)
    assert torch.equal(getattr(lhw, f"device_to_host_{sfx}")(r), torch.abs(a))
    getattr(lhw, f"set_const_val_{sfx}")(r, 7)
    assert torch.equal(getattr(lhw, f"device_to_host_{sfx}")(r), torch.full((6, 40), 7, dtype=dtype))


def test_moveaxis_and_pad(lhw, engine):
    print("[ext] moveaxis/pad", flush=True)
    a = torch.arange(2 * 3 * 4 * 5, dtype=torch.int64).reshape(2, 3, 4, 5)
    t = lhw.host_to_device_64(a)
    for s, d in [(0, 3), (3, 0), (1, -1), (-2, 1)]:
        m = engine.moveaxis(t, s, d)
        exp = torch.movedim(a, s, d)
        assert list(m.shape) == list(exp.shape) and m.is_contiguous()
        assert torch.equal(lhw.device_to_host_64(m), exp)
    for axis, before, after in [(0, 1, 2), (2, 0, 3), (-1, 4, 0)]:
        pdd = engine.pad_single_axis(t, axis, before, after)
        ax = axis % 4
        zb = list(a.shape)
        zb[ax] = before
        za = list(a.shape)
        za[ax] = after
        exp = torch.cat([torch.zeros(zb, dtype=torch.int64), a, torch.zeros(za, dtype=torch.int64)], dim=ax)
        assert torch.equal(lhw.device_to_host_64(pdd), exp)
    # a strided source (a transposed upload keeps its strides)
    at = a.permute(3, 1, 2, 0)
    m = engine.moveaxis(lhw.host_to_device_64(at), 0, 2)
    assert torch.equal(lhw.device_to_host_64(m), torch.movedim(at, 0, 2))
    with pytest.raises(ValueError):
        engine.moveaxis(t, 4, 0)
    with pytest.raises(ValueError):
        engine.pad_single_axis(t, 0, -1, 0)


def test_get_slice_views(lhw, engine):
    print("[ext] get_slice", flush=True)
    a = torch.arange(4 * 6 * 5, dtype=torch.int64).reshape(4, 6, 5)
    t = lhw.host_to_device_64(a)
    for key in [(1,), (slice(1, 3),), (slice(None), slice(1, 6, 2)), (2, slice(None, None, 3), 4), (Ellipsis, 1),
                (slice(-3, None), Ellipsis, slice(0, 5, 4)), (-1, -2), (slice(3, 1),)]:
        v = engine.get_slice(t, key)
        exp = a[key]
        assert list(v.shape) == list(exp.shape), key
        assert torch.equal(lhw.device_to_host_64(v), exp), key
    # a view writes through to its parent buffer
    v = engine.get_slice(t, (slice(1, 3), 2))
    lhw.set_const_val_64(v, -1)
    exp = a.clone()
    exp[1:3, 2] = -1
    assert torch.equal(lhw.device_to_host_64(t), exp)
    # and takes part in compute ops at its offset
    p = lhw.host_to_device_64(torch.tensor([7, 11, 13, 17, 19], dtype=torch.int64))
    row = engine.get_slice(t, (3, 4))
    r = lhw.allocate_on_hardware_64([5])
    lhw.modmul_ttt_64(row, row, p, r)
    x = exp[3, 4]
    assert torch.equal(lhw.device_to_host_64(r), (x * x) % torch.tensor([7, 11, 13, 17, 19]))
    with pytest.raises(IndexError):
        engine.get_slice(t, (4,))
    with pytest.raises(ValueError):
        engine.get_slice(t, (slice(None, None, -1),))


def test_butterfly_rate_hook(capi):
    """hx_butterfly_rate (the bench's VALU ceiling): a positive rate between a
    tenth and ten times 1024 SIMDs x 2.4 GHz / (100 cycles per wave of 64
    butterflies), repeatable within a factor of two."""
    nominal = 1024 * 2.4e9 * 64 / 100
    r1, r2 = capi.butterfly_rate(), capi.butterfly_rate()
    for r in (r1, r2):
        assert nominal / 10 < r < nominal * 10
    assert 0.5 < r1 / r2 < 2.0
