"""Transcripts replayed on the GPU through the runtime, the dispatcher and
lattica_hw, with exact verification of every device_to_host."""
from __future__ import annotations

import pytest
import torch

from heal_amd import params as P
from heal_amd.runtime import VerificationError, load_transcript_from_json, run_transcript, save_transcript_to_json
from transcript_builder import Builder

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine(gpu):
    from heal_amd.runtime.dispatcher import HealDispatcher

    return HealDispatcher()


def _plain_small_transcript(seed=0, corrupt=False):
    """Config: N = 4096, 4 primes of 30 bits (q = 1 mod 8192), 8 polynomials
    stored limb-major [8, 4, 4096] and passed as [8, 4096, 1, 4] views:
    ntt -> intt -> modmul with a second operand; expected = a * b mod q."""
    n = 4096
    primes = P.primes_below(30, 2 * n, 4)
    prm = P.NttParams(n, primes)
    g = torch.Generator().manual_seed(seed)
    q = torch.tensor(primes, dtype=torch.int64).view(1, 4, 1)
    a = torch.randint(0, 1 << 62, (8, 4, n), generator=g, dtype=torch.int64) % q
    b = torch.randint(0, 1 << 62, (8, 4, n), generator=g, dtype=torch.int64) % q
    a_view = a.permute(0, 2, 1).unsqueeze(2)  # [8, 4096, 1, 4], strides (16384, 1, ., 4096)
    b_view = b.permute(0, 2, 1)               # [8, 4096, 4]
    expected = (a.permute(0, 2, 1) * b_view) % torch.tensor(primes, dtype=torch.int64)
    if corrupt:
        expected = expected.clone()
        expected[3, 17, 2] += 1
    B = Builder()
    B.segment("plain")
    x = B.h2d(a_view)
    bb = B.h2d(b_view)
    p = B.h2d(prm.p)
    perm = B.h2d(prm.perm)
    tw = B.h2d(prm.twiddles)
    itw = B.h2d(prm.inv_twiddles)
    minv = B.h2d(prm.m_inv)
    f = B.empty([8, n, 1, 4])
    B.op("ntt", x, perm, None, p, None, None, tw, f, False, False, out=f)
    r = B.empty([8, n, 1, 4])
    B.op("intt", f, p, perm, itw, minv, None, None, r, out=r)
    B.op("reshape", r, [8, n, 4], out=r)
    o = B.empty([8, n, 4])
    B.op("_modmul_ttt", r, bb, p, o, out=o)
    B.d2h(o, expected)
    B.free(x)
    B.end()
    return B.transcript()


def test_plain_small_config_transcript(engine, tmp_path):
    path = tmp_path / "plain.json"
    save_transcript_to_json(_plain_small_transcript(), str(path))
    tr = load_transcript_from_json(str(path))
    stats = run_transcript(engine, tr.transcript, verify=True, timing=True, sync=torch.cuda.synchronize)
    assert stats["ops"] > 0 and stats["segments"][0][0] == "plain"


def test_verification_catches_a_wrong_value(engine):
    tr = _plain_small_transcript(seed=1, corrupt=True)
    with pytest.raises(VerificationError):
        run_transcript(engine, tr.transcript, verify=True)


def test_reference_dispatch_conventions(engine):
    """axis_modsum(a, axis, q, out), expand(a, repeat, axis), ntt tile flag, decomposition and permutation ops."""
    B = Builder()
    a = torch.tensor([[[1, 2], [3, 4], [5, 6]]], dtype=torch.int64)  # [1, 3, 2]
    q = torch.tensor([5, 7], dtype=torch.int64)
    x = B.h2d(a)
    qq = B.h2d(q)
    o = B.empty([1, 2])
    B.op("axis_modsum", x, 1, qq, o, out=o)
    B.d2h(o, a.sum(1) % q)
    col = B.h2d(torch.tensor([[3], [4]], dtype=torch.int64))
    B.op("expand", col, 3, 1, out=col)  # repeat=3 along axis 1
    e = B.empty([2, 3])
    B.op("_modsum_tcc", col, 0, 100, e, out=e)
    B.d2h(e, torch.tensor([[3, 3, 3], [4, 4, 4]]))
    v = B.h2d(torch.tensor([51, 29, 63], dtype=torch.int64))
    d = B.empty([3, 6])
    B.op("apply_g_decomp", v, d, 6, 1, out=d)
    B.d2h(d, torch.tensor([[(x >> i) & 1 for i in range(6)] for x in (51, 29, 63)]))
    t = B.h2d(torch.arange(12, dtype=torch.int64).reshape(1, 4, 3))
    pm = B.h2d(torch.tensor([[3, 2, 1, 0]] * 3, dtype=torch.int64))
    po = B.empty([1, 4, 3])
    B.op("take_along_axis", t, pm, po, 2, 1, out=po)
    B.d2h(po, torch.arange(12).reshape(1, 4, 3).flip(1))
    r = B.empty([3], torch.int64)
    B.op("_mod_tc", v, 10, r, out=r)
    B.d2h(r, torch.tensor([1, 9, 3]))
    run_transcript(engine, B.transcript().transcript, verify=True)
    with pytest.raises(NotImplementedError):
        engine.no_such_op(None, 0, 1)
