"""Benchmark: hybrid key-switch throughput at N = 2^16, L = 24, dnum = 3, w = 60
on one MI355X, with NTT limb-GB/s and the HBM roofline of the NTT pass kernel.

One step = the fused key-switch (three kernels launched on one stream) over a
batch of B independent key-switches (own x and own evaluation key each).  Timed
beside it: the hipGraph replay of the same kernels, the six HEAL calls over the
batch (eager and as a hipGraph) and the plain per-call form (B = 1); all must
agree bit for bit, and instance 0 must equal the oracle's chain.  Inputs are
resident in HBM before timing.  The cpu_baseline leg times the reference's CPU
implementation (compiled from its sources into oracle/_ref) on one key-switch
of the same shape and checks its output against the GPU's.

  python bench.py [--steps K] [--warmup W] [--batch B] [--gpus 1]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import re
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
SIMDS, CLOCK_HZ = 1024, 2.4e9  # MI355X: 256 CUs x 4 SIMDs at 2.4 GHz
# Profiled constants of the bench workload's kernels, keyed by (N, L, B); each
# entry names the profile it comes from.  HBM bytes per launch: rocprofv3
# --pmc FETCH_SIZE and WRITE_SIZE (one pass each), scaled by the same counter's
# reading for a plain device copy (FETCH_SIZE x 2, WRITE_SIZE x 1).  VALU
# instructions per launch: SQ_INSTS_VALU.  The NTT pass figures average its two
# launches.  "kernels" is the fingerprint of the kernel sources they were
# measured on (kernel_fingerprint()); on other sources the bench reports them
# as stale and leaves traffic and instruction counts null.
PROFILED = {
    (65536, 24, 16): {
        "kernels": "600ddb1e33776d1b",
        "ntt_pass_traffic": (1298222348, "profiles/ntt_pass_hbm_traffic.md"),
        "ntt_pass_valu": (128507904, "profiles/ntt_pass_sq_counters.md"),
        "k2_traffic": (2427036070, "profiles/ntt_pass_hbm_traffic.md"),
        "k2_valu": (252641280, "profiles/ntt_pass_sq_counters.md"),
    }
}
KERNEL_SOURCES = ("heal_amd/csrc/hx_arith.hpp", "heal_amd/csrc/hx_internal.hpp", "heal_amd/csrc/hx_ntt_kernels.hpp",
                  "heal_amd/csrc/hx_ntt.hip", "heal_amd/csrc/hx_keyswitch.hip")


def source_code_only(text: str) -> str:
    """C/C++ source without comments, whitespace runs collapsed."""
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    return re.sub(r"\s+", " ", text).strip()


def kernel_fingerprint() -> str:
    """SHA-256 (first 16 hex digits) of the kernel sources' code: a comment or
    layout edit keeps it, a code change does not."""
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(HERE, rel)) as f:
            h.update(source_code_only(f.read()).encode())
    return h.hexdigest()[:16]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--limbs", type=int, default=24)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-samples", type=int, default=5)  # about 10 s of host work
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 256 MiB device copy first (scales HBM byte counters under rocprofv3 --pmc)")
    return ap.parse_args()


def timed(fn, reps, torch):
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(reps):
        fn()
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / 1000.0


def cpu_baseline_leg(ks_gpu, samples):
    """Times the reference key-switch (oracle/_ref) on instance 0 of the batch."""
    import ctypes

    import torch

    sys.path.insert(0, os.path.join(HERE, "tests"))
    from oracle_lib import hview, reference

    ref = reference()
    if ref is None:
        return None, None
    N, L, D = ks_gpu.n, ks_gpu.L, ks_gpu.dnum
    prm = ks_gpu.prm
    x = ks_gpu.x[0].cpu()
    evk = ks_gpu.evk[:, 0].cpu().contiguous()
    host = {k: getattr(prm, k).cpu() for k in ("p", "perm", "twiddles", "inv_twiddles", "m_inv")}
    perms = ks_gpu.perms.cpu()
    out = torch.zeros((2, N, L), dtype=torch.int64)
    times = []
    for _ in range(samples):
        sec = ctypes.c_double(0)
        st = ref.ref_keyswitch(hview(x), hview(evk), hview(host["p"]), hview(host["perm"]), hview(host["twiddles"]),
                               hview(host["inv_twiddles"]), hview(host["m_inv"]), hview(perms), D, ks_gpu.w,
                               hview(out), ctypes.byref(sec))
        if st != 0:
            raise RuntimeError("reference key-switch failed: " + ref.ref_last_error().decode())
        times.append(sec.value)
    times.sort()
    med = times[len(times) // 2]
    # bit-exactness of the GPU output against the reference (valid where the
    # reference is defined: its INTT scaling overflows int64 for q > ~2^31.5,
    # so the comparison uses the oracle's exact chain instead when it differs)
    gpu_out = ks_gpu.out[0].cpu()
    return med, bool(torch.equal(out, gpu_out))


def oracle_chain_equal(ks_gpu):
    """The oracle's exact chain on instance 0 (the checker; not timed)."""
    import torch

    sys.path.insert(0, os.path.join(HERE, "tests"))
    from keyswitch_oracle import oracle_keyswitch

    ref_out = oracle_keyswitch(ks_gpu, 0)
    return bool(torch.equal(ref_out, ks_gpu.out[0].cpu()))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path shards by limb with no "
              "exchange; see the design notes)", file=sys.stderr)
        sys.exit(2)
    import torch

    from heal_amd.keyswitch import KeySwitchBatch

    torch.cuda.init()
    if args.copy_calibration:
        src = torch.ones(32 << 20, dtype=torch.int64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    N, L, B = args.n, args.limbs, args.batch
    batched = KeySwitchBatch(n=N, limbs=L, batch=B, seed=0)
    plain = KeySwitchBatch(n=N, limbs=L, batch=1, seed=0, primes=batched.primes)
    plain.x.copy_(batched.x[:1])
    plain.evk.copy_(batched.evk[:, :1])

    # correctness gate, bit for bit: plain chain == batched chain == chain graph
    # == fused kernels == fused graph
    plain.run()
    batched.run()
    torch.cuda.synchronize()
    ref_all = batched.out.clone()
    assert torch.equal(plain.out[0], ref_all[0]), "plain and batched forms differ"
    fused_out = torch.zeros_like(batched.out)
    batched.run_fused(fused_out)
    torch.cuda.synchronize()
    assert torch.equal(fused_out, ref_all), "fused form differs from the six-call chain"
    chain_graph = batched.capture(fused=False)
    batched.out.zero_()
    chain_graph.replay()
    torch.cuda.synchronize()
    assert torch.equal(batched.out, ref_all), "chain graph replay differs from eager"
    fused_graph = batched.capture(fused=True)
    batched.out.zero_()
    fused_graph.replay()
    torch.cuda.synchronize()
    assert torch.equal(batched.out, ref_all), "fused graph replay differs from eager"

    for _ in range(args.warmup):
        fused_graph.replay()
        chain_graph.replay()
        plain.run()
    torch.cuda.synchronize()

    t_graph = timed(fused_graph.replay, args.steps, torch)
    t_fused = timed(lambda: batched.run_fused(fused_out), args.steps, torch)
    # beside it: the batch with one evaluation key for every instance, as the
    # serving loop applies one rotation key to a batch (K2 then reads each key
    # block once per limb group, from L2 for the other instances); instance 0
    # has the same x and key as in the headline form
    key1 = batched.fused.prepare_key(batched.evk[:, :1].contiguous())
    shared_out = torch.empty_like(fused_out)
    t_shared = timed(lambda: batched.fused.exec_ex(batched.x, key1, shared_out, key_limbs=L, key_count=1),
                     args.steps, torch)
    torch.cuda.synchronize()
    assert torch.equal(shared_out[0], fused_out[0]), "shared-key form differs on instance 0"
    # the same K steps again with events between the three kernels (HIP events
    # on the launch stream), for the per-kernel times of the dominant kernel
    marks = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    for row in marks:
        for e in row:
            e.record()  # creates the underlying event
    torch.cuda.synchronize()
    for row in marks:
        batched.fused.exec_marked(batched.x, batched.key, fused_out, row)
    torch.cuda.synchronize()
    k_ms = [sum(row[i].elapsed_time(row[i + 1]) for row in marks) / args.steps for i in range(3)]
    t_marked = sum(marks[i][0].elapsed_time(marks[i][3]) for i in range(args.steps)) / 1e3
    t_chain_graph = timed(chain_graph.replay, args.steps, torch)
    t_batched = timed(batched.run, args.steps, torch)
    t_plain = timed(plain.run, args.steps * 2, torch) / 2.0
    ks_graph = args.steps * B / t_graph
    ks_fused = args.steps * B / t_fused
    ks_chain_graph = args.steps * B / t_chain_graph
    ks_batched = args.steps * B / t_batched
    ks_plain = args.steps / t_plain
    fused_gbs = batched.fused_bytes() * args.steps / t_fused / 1e9

    # NTT roofline: the forward digit NTTs of the batch (B*dnum*L limbs of N
    # points) run as two launches of the radix-16 pass kernel (one when N <= 2048).
    # Each launch runs half of every limb's stages, so its algorithmic share is
    # half of the transform's 2*N*8 bytes per limb: the re-read of the
    # intermediate between the launches is not credited.
    limbs_fwd = B * batched.dnum * L
    ntt_bytes = limbs_fwd * 2 * N * 8
    reps = 10
    dig_view = batched.digits.permute(0, 1, 3, 2)
    t_ntt = timed(lambda: batched.fwd(dig_view, batched.dntt), reps, torch)
    n_launch = 2 if N > 2048 else 1
    bytes_per_launch = ntt_bytes // n_launch
    avg_launch = t_ntt / (reps * n_launch)
    achieved = bytes_per_launch / avg_launch / 1e9
    ntt_limb_gbs = ntt_bytes * reps / t_ntt / 1e9
    limbs_inv = B * 2 * L
    t_intt = timed(lambda: batched.inv(batched.acc.view(B * 2, N, 1, L), batched.coef), reps, torch)
    intt_limb_gbs = limbs_inv * 2 * N * 8 * reps / t_intt / 1e9

    # HBM rates of the chain's elementwise calls on the batch, each timed alone:
    # algorithmic bytes = every operand read once and the result written once
    # (P = one [N, L] int64 polynomial per instance; the Galois table is read once)
    from heal_amd import capi
    D, P_ = batched.dnum, 8 * N * L * B
    dv = batched.dntt.permute(2, 0, 1, 3).unsqueeze(2)
    coef3, out3 = batched.coef.view(B * 2, N, L), batched.out.view(B * 2, N, L)
    p_ = batched.prm.p
    elem_ops = {
        "g_decomposition": (lambda: capi.g_decomposition(batched.x, batched.digits, D, batched.w), (1 + D) * P_),
        "modmul_ttt": (lambda: capi.modmul_ttt(dv, batched.evk, p_, batched.prod), 5 * D * P_),
        "axis_modsum": (lambda: capi.axis_modsum(batched.prod, p_, batched.acc, 0), (2 * D + 2) * P_),
        "permute": (lambda: capi.permute(coef3, batched.perms, out3, 2, 1, trusted=True, rows_same=batched.perms_same),
                    4 * P_ + batched.perms.numel() * batched.perms.element_size()),
    }
    elem_rates = {}
    for name, (fn, nbytes) in elem_ops.items():
        fn()
        t_op = timed(fn, reps, torch)
        gbs = nbytes * reps / t_op / 1e9
        elem_rates[name] = {"GBps": round(gbs, 1), "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 3),
                            "bytes": nbytes, "ms": round(t_op / reps * 1e3, 4)}
    torch.cuda.synchronize()
    assert torch.equal(batched.out, ref_all), "elementwise timing loop changed the chain's output"
    assert torch.equal(fused_out, ref_all), "the timed fused runs changed their output"

    prof = PROFILED.get((N, L, B), {})
    fingerprint = kernel_fingerprint()
    prof_current = prof.get("kernels") == fingerprint

    def profiled(key):
        v = prof.get(key)
        if not v:
            return None, None
        if not prof_current:
            return None, f"{v[1]} (measured on kernels {prof.get('kernels')}, these are {fingerprint}: stale, {v[0]})"
        return v[0], v[1]

    # K2: forward low window of the digit columns, key inner product, inverse
    # low window.  Its bytes as a kernel: the D digit-NTT intermediate columns
    # and the 2D key columns read, the 2 output columns written, per limb of
    # every key-switch: (3D + 2) * N * 8 B per limb.
    k2_bytes = (3 * D + 2) * N * 8 * L * B
    k2_s = k_ms[1] / 1e3
    k2_valu, k2_valu_src = profiled("k2_valu")
    k2_traffic, k2_traffic_src = profiled("k2_traffic")
    ntt_valu, ntt_valu_src = profiled("ntt_pass_valu")

    # VALU ceiling: the launch's radix-2 butterflies at the rate the GPU runs the
    # passes' own 60-bit lazy butterfly with its operands in registers
    # (hx_butterfly_rate, measured here); frac = that time / the launch time
    bfly_rate = capi.butterfly_rate()

    def valu_entry(insts, src, seconds, butterflies):
        e = {"butterflies_per_launch": butterflies, "butterfly_rate_G_per_s": round(bfly_rate / 1e9, 1),
             "ceiling_ms": round(butterflies / bfly_rate * 1e3, 4),
             "frac": round(butterflies / bfly_rate / seconds, 3)}
        if insts is not None:
            issue = insts * 4 / SIMDS / CLOCK_HZ  # quad-cycles on 1024 SIMDs
            e.update({"insts_per_launch": insts, "issue_ms_at_one_quad_cycle": round(issue * 1e3, 4),
                      "insts_source": src})
        return e

    # butterflies: a launch of the NTT pass runs log2(N)/2 stages of N/2 per limb;
    # K2 runs the inner window (8 stages of N/2) forward on D and inverse on 2
    # columns per limb of every instance
    logn = N.bit_length() - 1
    ntt_bflies = limbs_fwd * (N // 2) * (logn // n_launch)
    k2_bflies = B * L * (D + 2) * (N // 2) * batched.fused.config()["low_bits"]

    result = {
        "metric": "key-switches/sec (N=2^16, L=24, dnum=3, w=60), batched fused form; its hipGraph replay, the "
                  "plain per-call form and NTT limb-GB/s beside",
        "value": round(ks_fused, 2),
        "unit": "key-switches/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_fused / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seeded uniform residues mod 24 primes of 50-60 bits; seeded evaluation keys)",
        "config": {"workload": f"hybrid key-switch N={N} L={L} dnum=3 w=60, Galois g=5 permute", "batch_per_step": B,
                   "form": "fused key-switch, three kernel launches per step on one stream, B independent instances",
                   "plain_ks_per_s": round(ks_plain, 2),
                   "batched_chain_eager_ks_per_s": round(ks_batched, 2),
                   "batched_chain_graph_ks_per_s": round(ks_chain_graph, 2),
                   "fused_graph_ks_per_s": round(ks_graph, 2),
                   "fused_shared_key_ks_per_s": round(args.steps * B / t_shared, 2),
                   "fused_GBps_at_kernel_boundaries": round(fused_gbs, 1),
                   "ntt_limb_GBps": round(ntt_limb_gbs, 1), "intt_limb_GBps": round(intt_limb_gbs, 1),
                   "chain_call_hbm": elem_rates,
                   "keyswitch_algorithmic_bytes": batched.algorithmic_bytes() // B,
                   "keyswitch_chain_bytes": batched.chain_bytes() // B,
                   "keyswitch_fused_bytes": batched.fused_bytes() // B,
                   "fused_tuned_shape": batched.fused.config()},
        "roofline": {"bound": "hbm",
                     "kernel": f"k_ks_inner<8> (K2, the dominant kernel of the fused key-switch): {B} key-switches "
                               f"of N={N}, L={L}, dnum={D}",
                     "achieved": round(k2_bytes / k2_s / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(k2_bytes / k2_s / 1e9 / HBM_PEAK_GBS, 4), "traffic": k2_traffic,
                     "traffic_source": k2_traffic_src,
                     "bytes_per_launch": k2_bytes, "avg_launch_ms": round(k_ms[1], 4),
                     "kernel_ms": {"K1": round(k_ms[0], 4), "K2": round(k_ms[1], 4), "K3": round(k_ms[2], 4)},
                     "marked_step_ms": round(t_marked / args.steps * 1e3, 4),
                     "valu": valu_entry(k2_valu, k2_valu_src, k2_s, k2_bflies),
                     "keyswitch_GBps": round(batched.algorithmic_bytes() * args.steps / t_fused / 1e9, 1),
                     "keyswitch_frac": round(batched.algorithmic_bytes() * args.steps / t_fused / 1e9 / HBM_PEAK_GBS,
                                             4)},
        "roofline_ntt_pass": {"bound": "hbm",
                              "kernel": f"k_ntt_pass_r16<8,false>: forward NTT of {limbs_fwd} limbs of N={N}, "
                                        f"{n_launch} launches per NTT",
                              "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": profiled("ntt_pass_traffic")[0],
                              "traffic_source": profiled("ntt_pass_traffic")[1],
                              "bytes_per_launch": bytes_per_launch, "avg_launch_ms": round(avg_launch * 1e3, 4),
                              "launch_read_write_GBps": round(2 * achieved, 1),
                              "valu": valu_entry(ntt_valu, ntt_valu_src, avg_launch, ntt_bflies)},
        "cpu_baseline": None,
    }
    try:
        result["config"]["oracle_bit_exact"] = oracle_chain_equal(plain)
    except Exception as e:  # the checker is optional on the box; never the measured path
        result["config"]["oracle_bit_exact"] = f"unchecked: {e}"
    if not args.no_cpu_baseline:
        try:
            med, equal = cpu_baseline_leg(plain, args.cpu_samples)
            if med is not None:
                cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
                result["cpu_baseline"] = {
                    "value": round(1.0 / med, 4), "unit": "key-switches/s", "cores": cores, "kind": "reference",
                    "sample": f"{args.cpu_samples} full-size key-switches through the reference example_impl "
                              f"(median of compute time, {cores} OpenMP threads)",
                    "gpu_over_cpu": round(ks_fused * med, 1),
                    "reference_output_equal": equal,
                }
        except Exception as e:
            result["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(result), flush=True)
    batched.close()
    plain.close()


if __name__ == "__main__":
    main()
