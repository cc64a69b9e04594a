"""The key-switch chain on the CPU oracle (test infrastructure).

Same six calls and views as heal_amd.keyswitch.KeySwitchBatch.run(), on
instance `b` of a batch, through oracle/heal_oracle.c.  mode=1 uses the
reference's wrapped INTT scaling (for comparison with oracle/_ref).
"""
from __future__ import annotations

import torch

from heal_cases import OracleBackend


def oracle_keyswitch(ks, b: int = 0, mode: int = 0):
    O = OracleBackend()
    N, L, D, w = ks.n, ks.L, ks.dnum, ks.w
    prm = ks.prm
    p, perm, tw, itw, minv = (getattr(prm, k).cpu() for k in ("p", "perm", "twiddles", "inv_twiddles", "m_inv"))
    x = ks.x[b].cpu()
    evk = ks.evk[:, b].cpu()
    perms = ks.perms.cpu()
    digits = torch.zeros((N, L, D), dtype=torch.int64)
    st, _ = O.g_decomposition(x, digits, D, w)
    assert st is None
    dntt = torch.zeros((1, N, D, L), dtype=torch.int64)
    assert O.ntt(digits.view(1, N, L, D).permute(0, 1, 3, 2), p, perm, tw, dntt) is None
    prod = torch.zeros((D, 2, N, L), dtype=torch.int64)
    assert O.modop(0, dntt.view(N, D, L).permute(1, 0, 2).unsqueeze(1), evk, 0, p, 0, prod) is None
    acc = torch.zeros((2, N, L), dtype=torch.int64)
    assert O.axis_modsum(prod, p, acc, 0) is None
    coef = torch.zeros((2, N, 1, L), dtype=torch.int64)
    assert O.intt(acc.view(2, N, 1, L), p, perm, itw, minv, coef, mode) is None
    out = torch.zeros((2, N, L), dtype=torch.int64)
    assert O.permute(coef.view(2, N, L), perms, out, 2, 1) is None
    return out
