"""TranscriptProgram: a transcript replayed many times, with its segment
captured as a hipGraph on the second run and replayed afterwards.  Every run
verifies its device_to_host results exactly; inputs change between runs."""
from __future__ import annotations

import pytest
import torch

from heal_amd import params as P
from transcript_builder import Builder

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine(gpu):
    from heal_amd.runtime.dispatcher import HealDispatcher

    return HealDispatcher()


def _program(n=4096, k=4):
    """x -> ntt -> intt -> modmul with b -> axis_modsum over the batch, inside one segment."""
    primes = P.primes_below(30, 2 * n, k)
    prm = P.NttParams(n, primes)
    B = Builder()
    x = B.h2d(torch.zeros(8, n, 1, k, dtype=torch.int64))
    h_x = B.ops[-1][1].args[0].value
    b = B.h2d(torch.zeros(8, n, k, dtype=torch.int64))
    h_b = B.ops[-1][1].args[0].value
    p = B.h2d(prm.p)
    h_p = B.ops[-1][1].args[0].value
    perm = B.h2d(prm.perm)
    h_perm = B.ops[-1][1].args[0].value
    tw = B.h2d(prm.twiddles)
    itw = B.h2d(prm.inv_twiddles)
    minv = B.h2d(prm.m_inv)
    B.segment("layer")
    f = B.empty([8, n, 1, k])
    B.op("ntt", x, perm, None, p, None, None, tw, f, False, False, out=f)
    r = B.empty([8, n, 1, k])
    B.op("intt", f, p, perm, itw, minv, None, None, r, out=r)
    B.op("reshape", r, [8, n, k], out=r)
    o = B.empty([8, n, k])
    B.op("_modmul_ttt", r, b, p, o, out=o)
    s = B.zeros([n, k])
    B.op("axis_modsum", o, 0, p, s, out=s)
    B.end()
    B.d2h(o, torch.zeros(8, n, k, dtype=torch.int64))
    h_o = B.ops[-1][1].out.value
    B.d2h(s, torch.zeros(n, k, dtype=torch.int64))
    h_s = B.ops[-1][1].out.value
    return B.transcript().transcript, primes, (h_x, h_b, h_perm, h_o, h_s)


def _set_run(handles, primes, seed, n=4096, k=4):
    h_x, h_b, _, h_o, h_s = handles
    g = torch.Generator().manual_seed(seed)
    q = torch.tensor(primes, dtype=torch.int64)
    xr = torch.randint(0, 1 << 62, (8, n, k), generator=g) % q
    br = torch.randint(0, 1 << 62, (8, n, k), generator=g) % q
    h_x.tensor = xr.view(8, n, 1, k).clone()
    h_b.tensor = br
    prod = torch.tensor([[[(int(a) * int(c)) % int(m) for a, c, m in zip(ra, rb, primes)]
                          for ra, rb in zip(xa, xb)] for xa, xb in zip(xr.tolist(), br.tolist())], dtype=torch.int64)
    h_o.tensor = prod
    h_s.tensor = prod.sum(0) % q


def test_segment_graph_capture_and_replay(engine):
    from heal_amd.runtime import TranscriptProgram, VerificationError

    n = 1024
    tr, primes, handles = _program(n=n)
    prog = TranscriptProgram(engine, tr)
    # run 1 eager (prepares plans), run 2 sees the same buffers and captures, then replays
    expect = ["eager", "capture", "replay", "replay"]
    for run, mode in enumerate(expect):
        print(f"[program] run {run}", flush=True)
        _set_run(handles, primes, 100 + run, n=n)
        stats = prog.run(verify=True)
        assert [m for _, m, _ in stats["segments"]] == [mode], (run, stats)
    # a corrupted expectation is still caught on a replayed run
    _set_run(handles, primes, 200, n=n)
    handles[3].tensor = handles[3].tensor.clone()
    handles[3].tensor[1, 2, 3] += 1
    with pytest.raises(VerificationError):
        prog.run(verify=True)
    # the permutation is a table: rewriting it drops the graph; the NTT/INTT
    # round trip does not depend on which permutation both use
    h_perm = handles[2]
    h_perm.tensor = torch.arange(n, dtype=torch.int64)
    modes = []
    for run in range(3):
        _set_run(handles, primes, 210 + run, n=n)
        modes += [m for _, m, _ in prog.run(verify=True)["segments"]]
    assert modes == ["eager", "capture", "replay"], modes


def test_eager_and_graph_runs_agree(engine):
    """The same inputs through a program without graphs and one with them."""
    from heal_amd.runtime import TranscriptProgram

    n = 2048
    tr, primes, handles = _program(n=n)
    eager = TranscriptProgram(engine, tr, graphs=False)
    graphed = TranscriptProgram(engine, tr, graphs=True)
    for run in range(3):
        _set_run(handles, primes, 300 + run, n=n)
        a = eager.run(verify=True)
        b = graphed.run(verify=True)
        assert a["graph_replays"] == 0 and a["captures"] == 0
    assert b["graph_replays"] == 1
    assert [m for _, m, _ in b["segments"]] == ["replay"]
