"""Plans shared by threads, streams and captured graphs (the C++ HEAL API,
through its pybind11 module `lattica_hw`).

* The same NTT tables on two streams from two threads: each stream gets its
  own workspace, so interleaved executions do not corrupt each other; both
  results equal the oracle.
* A graph captured with NTT plans keeps them (capture_begin / capture_end):
  after more distinct plan keys than the plan cache holds, replays still give
  the oracle's result; dropping the record releases the plans.
* Plan destruction and status polls do not wait for other streams."""
from __future__ import annotations

import threading

import pytest
import torch

from heal_cases import OracleBackend
from heal_amd import params as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lhw(gpu):
    from heal_amd import lattica_hw

    return lattica_hw


def _tables(lhw, m, L):
    primes = P.keyswitch_primes(L)
    prm = P.NttParams(m, primes)
    dev = {k: lhw.host_to_device_64(getattr(prm, k)) for k in ("p", "perm", "twiddles", "inv_twiddles", "m_inv")}
    return prm, primes, dev


def test_same_tables_on_two_streams_from_two_threads(lhw):
    m, L, l = 16384, 8, 4
    prm, primes, T = _tables(lhw, m, L)
    g = torch.Generator().manual_seed(31)
    inputs = [torch.stack([torch.randint(0, q, (l, m, 1), generator=g) for q in primes], dim=-1) for _ in range(2)]
    want = []
    for a in inputs:
        ro = torch.zeros_like(a)
        assert OracleBackend().ntt(a, prm.p, prm.perm, prm.twiddles, ro) is None
        want.append(ro)
    results = [None, None]
    errors = []
    barrier = threading.Barrier(2)

    def worker(i):
        try:
            s = torch.cuda.Stream()
            lhw.set_stream(s.cuda_stream)
            a = lhw.host_to_device_64(inputs[i])
            r = lhw.allocate_on_hardware_64(list(inputs[i].shape))
            barrier.wait()
            for _ in range(8):  # many executions in flight on both streams
                lhw.ntt_64(a, T["p"], T["perm"], T["twiddles"], None, None, r)
            results[i] = lhw.device_to_host_64(r)
            # the inverse on the same stream returns the input
            back = lhw.allocate_on_hardware_64(list(inputs[i].shape))
            lhw.intt_64(r, T["p"], T["perm"], T["inv_twiddles"], T["m_inv"], None, None, back)
            assert torch.equal(lhw.device_to_host_64(back), inputs[i])
        except Exception as e:  # reported by the main thread
            errors.append(e)
        finally:
            lhw.set_stream(0)

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for i in range(2):
        assert torch.equal(results[i], want[i]), i


def test_capture_record_keeps_plans_past_cache_eviction(lhw):
    """A captured NTT keeps its plan after many other tables pushed it out of
    the plan cache; replays still equal the oracle."""
    m, L = 4096, 4
    prm, primes, T = _tables(lhw, m, L)
    g = torch.Generator().manual_seed(5)
    a_host = torch.stack([torch.randint(0, q, (2, m, 1), generator=g) for q in primes], dim=-1)
    want = torch.zeros_like(a_host)
    assert OracleBackend().ntt(a_host, prm.p, prm.perm, prm.twiddles, want) is None
    a = lhw.host_to_device_64(a_host)
    r = lhw.allocate_on_hardware_64(list(a_host.shape))
    s = torch.cuda.Stream()
    lhw.set_stream(s.cuda_stream)
    try:
        lhw.ntt_64(a, T["p"], T["perm"], T["twiddles"], None, None, r)  # plan created outside capture
        torch.cuda.synchronize()
        before = lhw.captured_plan_count()
        graph = torch.cuda.CUDAGraph()
        lhw.capture_begin()
        try:
            with torch.cuda.graph(graph, stream=s):
                lhw.ntt_64(a, T["p"], T["perm"], T["twiddles"], None, None, r)
        finally:
            pins = lhw.capture_end()
        assert lhw.captured_plan_count() == before + 1
        # more distinct plan keys than the cache keeps: each re-upload is a new buffer
        for i in range(80):
            tw_i = lhw.host_to_device_64(prm.twiddles)
            scratch = lhw.allocate_on_hardware_64(list(a_host.shape))
            lhw.ntt_64(a, T["p"], T["perm"], tw_i, None, None, scratch)
        torch.cuda.synchronize()
        for rep in range(3):
            lhw.set_const_val_64(r, 0)
            torch.cuda.synchronize()
            graph.replay()
            torch.cuda.synchronize()
            assert torch.equal(lhw.device_to_host_64(r), want), rep
        torch.cuda.synchronize()
        del graph
        del pins
        assert lhw.captured_plan_count() == before
    finally:
        lhw.set_stream(0)


def test_plan_destroy_and_poll_do_not_wait_for_other_streams(gpu):
    """Destroying a plan and polling the status while another stream has a long
    queue of work return at once (memory is released in stream order, the poll
    follows the polling stream only), and the queued work still completes
    correctly."""
    import time

    from heal_amd import capi

    m, L = 65536, 24
    prm = P.NttParams(m, P.keyswitch_primes(L)).to("cuda")
    busy = torch.cuda.Stream()
    x = torch.randint(0, 1 << 40, (16, m, 1, L), dtype=torch.int64, device="cuda") % prm.p
    y = torch.empty_like(x)
    fwd = capi.NttPlan(prm.p, prm.perm, prm.twiddles)
    fwd(x, y)
    torch.cuda.synchronize()
    want = y.clone()
    other = capi.NttPlan(prm.p, prm.perm, prm.twiddles)
    torch.cuda.synchronize()
    with torch.cuda.stream(busy):
        for _ in range(400):  # a few hundred milliseconds of queued work
            fwd(x, y)
    t0 = time.perf_counter()
    other.close()
    capi.status_poll()  # the current stream only
    dt = time.perf_counter() - t0
    busy_left = not torch.cuda.Stream.query(busy)
    torch.cuda.synchronize()
    assert busy_left, "the queued work finished before the measurement; the check is void"
    assert dt < 0.05, dt
    assert torch.equal(y, want)
    fwd.close()
