"""Child process of test_gpu_keyswitch_tune.py (not a test module): creates
fused key-switch plans under the tuning cache named by HEAL_AMD_TUNE_CACHE
(read once per process, hence a process of its own) and prints one JSON line
per check.

mode "tune":  one plan at N = 4096, L = 3, dnum = 3; prints its configuration.
mode "check": plans at N = 4096, 65536 and 32768; prints their configurations and
              whether the fused key-switch equals the six-call chain, whether
              hx_keyswitch_exec_rmac equals hx_keyswitch_exec_ex followed by
              hx_rot_mac, and whether a group's members (32-column tiles)
              equal their own calls."""
from __future__ import annotations

import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import torch  # noqa: E402

from heal_amd import capi  # noqa: E402
from heal_amd import params as P  # noqa: E402
from heal_amd.keyswitch import KeySwitchBatch, UnevenKeySwitch  # noqa: E402


def plans(n, L, dnum=3, w=60):
    prm = P.NttParams(n, P.keyswitch_primes(L), device="cuda")
    fwd = capi.NttPlan(prm.p, prm.perm, prm.twiddles)
    inv = capi.NttPlan(prm.p, prm.perm, prm.inv_twiddles, prm.m_inv, inverse=True)
    ks = capi.KeySwitchPlan(fwd, inv, P.galois_perms(n, 5, L, device="cuda"), dnum, w)
    return prm, fwd, inv, ks


def tune():
    _, fwd, inv, ks = plans(4096, 3)
    print(json.dumps({"n": 4096, "L": 3, "config": ks.config()}), flush=True)
    for p in (ks, fwd, inv):
        p.close()


def fused_vs_chain(n, L, batch, seed):
    ks = KeySwitchBatch(n=n, limbs=L, dnum=3, base_bits=20, batch=batch, seed=seed)
    ks.run()
    torch.cuda.synchronize()
    want = ks.out.clone()
    out = torch.zeros_like(ks.out)
    ks.run_fused(out)
    torch.cuda.synchronize()
    res = {"check": "fused", "n": n, "L": L, "config": ks.fused.config(), "equal": bool(torch.equal(out, want))}
    ks.close()
    return res


def rmac_vs_separate(n, L, B, seed, M=128, D=3):
    prm, fwd, inv, ks = plans(n, L, D, 60)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    q = prm.p
    ct = torch.randint(0, 1 << 62, (B, 2, n, L), generator=g, device="cuda") % q
    evk = torch.randint(0, 1 << 62, (D, 1, 2, n, L), generator=g, device="cuda") % q
    key = ks.prepare_key(evk)
    acc0 = torch.randint(-(1 << 62), 1 << 62, (B, 2, n, L), generator=g, device="cuda")
    diags = torch.randint(-(1 << 62), 1 << 62, (B, M, M, L), generator=g, device="cuda")
    perm = P.galois_perms(n, pow(5, 7, 2 * n), 1, device="cuda")[0].contiguous()
    rmap = ks.row_map(perm)
    out = torch.empty_like(ct)
    ks.exec_ex(ct[:, 1], key, out, key_limbs=L, key_count=1, row_map=rmap)
    want = acc0.clone()
    capi.rot_mac(want, out, ct, perm, diags, 7, q)
    got = acc0.clone()
    ks.exec_rmac(ct[:, 1], key, got, ct, diags, 7, key_limbs=L, key_count=1, row_map=rmap)
    torch.cuda.synchronize()
    res = {"check": "rmac", "n": n, "L": L, "config": ks.config(), "equal": bool(torch.equal(got, want))}
    for p in (ks, fwd, inv):
        p.close()
    return res


def group_vs_members(seed):
    u = UnevenKeySwitch(count=5, ns=(4096,), limbs=(2, 4), seed=seed)
    outs = [torch.full_like(o, -1) for o in u.outs]
    grp = u.group(outs)
    grp()
    u.run_plain()
    torch.cuda.synchronize()
    equal = all(torch.equal(outs[i], u.outs[i]) for i in range(len(outs)))
    tiles = sorted({plan.config()["tile"] for plan, _, _, _ in u.plans.values()})
    grp.close()
    return {"check": "group", "tiles": tiles, "equal": bool(equal)}


def check():
    for res in (fused_vs_chain(4096, 3, 2, 3), fused_vs_chain(65536, 24, 2, 5), fused_vs_chain(32768, 16, 2, 6),
                rmac_vs_separate(4096, 4, 3, 21), rmac_vs_separate(65536, 24, 2, 22),
                rmac_vs_separate(32768, 16, 2, 23), group_vs_members(7)):
        print(json.dumps(res), flush=True)


if __name__ == "__main__":
    {"tune": tune, "check": check}[sys.argv[1]]()
