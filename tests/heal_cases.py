"""Runs golden cases (tests/golden/reference_kat.json) against a backend.

A backend exposes the HEAL operations over torch tensors on its device and
returns a status name (None on success, else "invalid_argument",
"out_of_range", ...).  Three backends exist: the C oracle and the compiled
reference (host memory, in tests/oracle_lib.py) and the MI355X C ABI
(device memory, GpuBackend below).
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np
import torch

from oracle_lib import CODES, hview, oracle, reference

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

_DT = {"int32": torch.int32, "int64": torch.int64, "float64": torch.float64}


def load_cases(name="reference_kat.json"):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def tensor(spec, device="cpu"):
    if spec is None:
        return None
    t = torch.tensor(np.array(spec["data"], dtype=np.int64).reshape(spec["shape"]), dtype=_DT[spec["dtype"]])
    t = t.to(device)
    for op in spec.get("view", []):
        if op[0] == "transpose":
            t = t.transpose(op[1], op[2])
        elif op[0] == "permute":
            t = t.permute(*op[1])
    return t


class OracleBackend:
    device = "cpu"

    def __init__(self, lib=None):
        self.o = lib or oracle()

    def modop(self, op, a, b, bs, p, ps, r):
        return CODES[self.o.or_modop(op, hview(a), hview(b), bs, hview(p), ps, hview(r))]

    def mod(self, a, as_, b, bs, r, a_t, b_t):
        return CODES[self.o.or_mod(hview(a), as_, hview(b), bs, hview(r), int(a_t), int(b_t))]

    def axis_modsum(self, a, p, r, axis):
        return CODES[self.o.or_axis_modsum(hview(a), hview(p), hview(r), axis)]

    def g_decomposition(self, a, r, power, bb):
        ov = ctypes.c_uint64(0)
        code = self.o.or_g_decomposition(hview(a), hview(r), power, bb, ctypes.byref(ov))
        return CODES[code], ov.value

    def ntt(self, a, p, perm, tw, r):
        return CODES[self.o.or_ntt(hview(a), hview(p), hview(perm), hview(tw), hview(r))]

    def intt(self, a, p, perm, itw, minv, r, mode=0):
        return CODES[self.o.or_intt(hview(a), hview(p), hview(perm), hview(itw), hview(minv), hview(r), mode)]

    def permute(self, a, perms, r, ew, pa):
        return CODES[self.o.or_permute(hview(a), hview(perms), hview(r), ew, pa)]

    def strided_copy(self, s, d):
        return CODES[self.o.or_strided_copy(hview(s), hview(d))]


class RefBackend:
    """The compiled reference (oracle/_ref), same calling convention."""

    device = "cpu"

    def __init__(self):
        self.r = reference()

    def modop(self, op, a, b, bs, p, ps, r):
        return CODES[self.r.ref_modop(op, hview(a), hview(b), bs, hview(p), ps, hview(r))]

    def mod(self, a, as_, b, bs, r, a_t, b_t):
        variant = 0 if (a_t and b_t) else (1 if a_t else 2)
        return CODES[self.r.ref_mod(variant, hview(a), as_, hview(b), bs, hview(r))]

    def axis_modsum(self, a, p, r, axis):
        return CODES[self.r.ref_axis_modsum(hview(a), hview(p), hview(r), axis)]

    def g_decomposition(self, a, r, power, bb):
        return CODES[self.r.ref_g_decomposition(hview(a), hview(r), power, bb)], None

    def ntt(self, a, p, perm, tw, r):
        return CODES[self.r.ref_ntt(hview(a), hview(p), hview(perm), hview(tw), hview(r))]

    def intt(self, a, p, perm, itw, minv, r, mode=0):
        return CODES[self.r.ref_intt(hview(a), hview(p), hview(perm), hview(itw), hview(minv), hview(r))]

    def permute(self, a, perms, r, ew, pa):
        return CODES[self.r.ref_permute(hview(a), hview(perms), hview(r), ew, pa)]

    def strided_copy(self, s, d):
        return CODES[self.r.ref_make_contiguous(hview(s), hview(d))]


def _exc_code(fn):
    try:
        fn()
        return None
    except ValueError:
        return "invalid_argument"
    except IndexError:
        return "out_of_range"
    except MemoryError:
        return "bad_alloc"


class GpuBackend:
    """The MI355X library through its C ABI (heal_amd.capi)."""

    device = "cuda"

    def __init__(self):
        from heal_amd import capi

        self.c = capi

    def modop(self, op, a, b, bs, p, ps, r):
        c = self.c
        fns = {(True, True): (c.modmul_ttt, c.modsum_ttt), (True, False): (c.modmul_ttc, c.modsum_ttc),
               (False, True): (c.modmul_tct, c.modsum_tct), (False, False): (c.modmul_tcc, c.modsum_tcc)}
        fn = fns[(b is not None, p is not None)][op]
        return _exc_code(lambda: fn(a, b if b is not None else bs, p if p is not None else ps, r))

    def mod(self, a, as_, b, bs, r, a_t, b_t):
        c = self.c
        if a_t and b_t:
            return _exc_code(lambda: c.check(c.lib.hx_mod_tt(c.view(a), c.view(b), c.view(r), c.stream_of(r))))
        if a_t:
            return _exc_code(lambda: c.check(c.lib.hx_mod_tc(c.view(a), bs, c.view(r), c.stream_of(r))))
        return _exc_code(lambda: c.check(c.lib.hx_mod_ct(as_, c.view(b), c.view(r), c.stream_of(r))))

    def axis_modsum(self, a, p, r, axis):
        return _exc_code(lambda: self.c.axis_modsum(a, p, r, axis))

    def g_decomposition(self, a, r, power, bb):
        self.c.status_poll()
        code = _exc_code(lambda: self.c.g_decomposition(a, r, power, bb))
        st = self.c.status_poll()
        return code, st["gdecomp_overflow"]

    def ntt(self, a, p, perm, tw, r):
        return _exc_code(lambda: self.c.ntt(a, p, perm, tw, r))

    def intt(self, a, p, perm, itw, minv, r, mode=0):
        return _exc_code(lambda: self.c.intt(a, p, perm, itw, minv, r))

    def permute(self, a, perms, r, ew, pa):
        return _exc_code(lambda: self.c.permute(a, perms, r, ew, pa))

    def strided_copy(self, s, d):
        return _exc_code(lambda: self.c.strided_copy(s, d))


class LatticaBackend:
    """The pybind11 module `lattica_hw` (C++ HEAL API templates over the C ABI).
    Takes host tensors, uploads them with host_to_device_<T> (strides kept),
    runs the op and downloads the result into the caller's tensor."""

    device = "cpu"
    _SUFFIX = {torch.int32: "32", torch.int64: "64", torch.float64: "float64"}

    def __init__(self):
        from heal_amd import lattica_hw

        self.m = lattica_hw

    def fn(self, name, t):
        return getattr(self.m, f"{name}_{self._SUFFIX[t.dtype]}")

    def up(self, t):
        return None if t is None else self.fn("host_to_device", t)(t)

    def down(self, hw, t):
        t.copy_(self.fn("device_to_host", t)(hw))

    def _run(self, body, out, out_hw):
        code = _exc_code(body)
        if code is None:
            self.down(out_hw, out)
        return code

    def modop(self, op, a, b, bs, p, ps, r):
        variant = ("t" if b is not None else "c") + ("t" if p is not None else "c")
        f = self.fn(("modmul_t" if op == 0 else "modsum_t") + variant, r)
        A, B, Pp, R = self.up(a), self.up(b), self.up(p), self.up(r)
        return self._run(lambda: f(A, B if b is not None else bs, Pp if p is not None else ps, R), r, R)

    def mod(self, a, as_, b, bs, r, a_t, b_t):
        A, B, R = self.up(a), self.up(b), self.up(r)
        if a_t and b_t:
            return self._run(lambda: self.fn("mod_tt", r)(A, B, R), r, R)
        if a_t:
            return self._run(lambda: self.fn("mod_tc", r)(A, bs, R), r, R)
        return self._run(lambda: self.fn("mod_ct", r)(as_, B, R), r, R)

    def axis_modsum(self, a, p, r, axis):
        A, Pp, R = self.up(a), self.up(p), self.up(r)
        return self._run(lambda: self.fn("axis_modsum", r)(A, Pp, R, axis), r, R)

    def g_decomposition(self, a, r, power, bb):
        A, R = self.up(a), self.up(r)
        return self._run(lambda: self.fn("g_decomposition", r)(A, R, power, bb), r, R), None

    def ntt(self, a, p, perm, tw, r):
        A, Pp, PM, TW, R = self.up(a), self.up(p), self.up(perm), self.up(tw), self.up(r)
        return self._run(lambda: self.fn("ntt", r)(A, Pp, PM, TW, None, None, R), r, R)

    def intt(self, a, p, perm, itw, minv, r, mode=0):
        A, Pp, PM, TW, MI, R = self.up(a), self.up(p), self.up(perm), self.up(itw), self.up(minv), self.up(r)
        return self._run(lambda: self.fn("intt", r)(A, Pp, PM, TW, MI, None, None, R), r, R)

    def permute(self, a, perms, r, ew, pa):
        A, PM, R = self.up(a), self.up(perms), self.up(r)
        return self._run(lambda: self.fn("permute", r)(A, PM, R, ew, pa), r, R)

    def strided_copy(self, s, d):
        S = self.up(s)
        code = _exc_code(lambda: self.fn("make_contiguous", s)(S))
        if code is None:
            self.down(S, d)
        return code

    # metadata-only view ops of the C++ API
    def view_op(self, case):
        A = case["args"]
        a = tensor(A["a"])
        hw = self.up(a)
        op = case["op"]
        try:
            if op == "squeeze":
                res = self.fn("squeeze", a)(hw, A["axis"])
            elif op == "unsqueeze":
                res = self.fn("unsqueeze", a)(hw, A["axis"])
            elif op == "reshape":
                hw.reshape(A["shape"])
                res = hw
            elif op == "make_contiguous":
                res = self.fn("make_contiguous", a)(hw)
                return None, self.fn("device_to_host", a)(res), {"same_object": res is hw}
            else:
                return "skip", None, {}
        except ValueError:
            return "invalid_argument", None, {}
        except IndexError:
            return "out_of_range", None, {}
        return None, self.fn("device_to_host", a)(res), {"same_object": res is hw}


def run_case(case, be):
    """Returns (status, result_tensor_on_cpu or None, extra dict)."""
    dev = be.device
    A = case["args"]
    op = case["op"]
    ex = {}

    def t(name):
        return tensor(A.get(name), dev)

    def out_like(shape, dtype):
        return torch.zeros(shape, dtype=_DT[dtype], device=dev)

    if op in ("modmul", "modsum"):
        a = t("a")
        b = A["b"]
        p = A["p"]
        bt = tensor(b, dev) if isinstance(b, dict) else None
        pt = tensor(p, dev) if isinstance(p, dict) else None
        r = out_like(A["result_shape"], A["a"]["dtype"])
        st = be.modop(0 if op == "modmul" else 1, a, bt, 0 if bt is not None else b, pt, 0 if pt is not None else p, r)
        return st, r.cpu(), ex
    if op == "mod":
        a, b = A["a"], A["b"]
        # scalars are ints; None is a null tensor in a tensor slot
        a_t = isinstance(a, dict) or a is None
        b_t = isinstance(b, dict) or b is None
        at = tensor(a, dev) if isinstance(a, dict) else None
        bt = tensor(b, dev) if isinstance(b, dict) else None
        dtype = (a if isinstance(a, dict) else b)["dtype"] if (isinstance(a, dict) or isinstance(b, dict)) else "int64"
        r = out_like(A["result_shape"], dtype)
        st = be.mod(at, a if isinstance(a, int) else 0, bt, b if isinstance(b, int) else 0, r, a_t, b_t)
        return st, r.cpu(), ex
    if op == "axis_modsum":
        a, p = t("a"), t("p")
        r = out_like(A["result_shape"], A["a"]["dtype"])
        return be.axis_modsum(a, p, r, A["axis"]), r.cpu(), ex
    if op == "g_decomposition":
        a = t("a")
        r = out_like(A["result_shape"], A["a"]["dtype"])
        st, ov = be.g_decomposition(a, r, A["power"], A["base_bits"])
        ex["overflow"] = ov
        return st, r.cpu(), ex
    if op == "ntt":
        a = t("a")
        r = torch.zeros_like(a)
        return be.ntt(a, t("p"), t("perm"), t("twiddles"), r), r.cpu(), ex
    if op == "intt":
        a = t("a")
        r = torch.zeros_like(a)
        return be.intt(a, t("p"), t("perm"), t("twiddles"), t("m_inv"), r), r.cpu(), ex
    if op == "permute":
        a = t("a")
        r = torch.zeros_like(a)
        return be.permute(a, t("perms"), r, A["elementwise_axis"], A["perm_axis"]), r.cpu(), ex
    if op == "decomp_reconstruct":
        a = t("a")
        n = a.shape[0]
        digits = torch.zeros((n, A["power"]), dtype=a.dtype, device=dev)
        st, _ = be.g_decomposition(a, digits, A["power"], A["base_bits"])
        if st:
            return st, None, ex
        st = be.modop(0, digits, t("basis"), 0, t("p6"), 0, digits)  # in place, as the reference test
        if st:
            return st, None, ex
        d3 = digits.view(n, A["power"], 1)
        rec = torch.zeros((n, 1), dtype=a.dtype, device=dev)
        st = be.axis_modsum(d3, t("p1"), rec, 1)
        return st, rec.view(n).cpu(), ex
    if op == "make_contiguous":
        a = t("a")
        r = torch.zeros(a.shape, dtype=a.dtype, device=dev)
        return be.strided_copy(a, r), r.cpu(), ex
    return "skip", None, ex


def check_case(case, be):
    if case["op"] in ("squeeze", "unsqueeze", "reshape") or (
            case["op"] == "make_contiguous" and hasattr(be, "view_op")):
        if not hasattr(be, "view_op"):
            return "skip"
        st, r, ex = be.view_op(case)
    else:
        st, r, ex = run_case(case, be)
    if st == "skip":
        return "skip"
    exp = case["expect"]
    if "error" in exp:
        assert st == exp["error"], f"{case['name']}: expected {exp['error']}, got {st}"
        return "ok"
    if "same_object" in exp:
        assert st is None
        assert torch.equal(r, tensor(case["args"]["a"]))
        if "same_object" in ex:
            assert ex["same_object"], f"{case['name']}: make_contiguous must return the same tensor"
        return "ok"
    assert st is None, f"{case['name']}: unexpected error {st}"
    want = tensor(exp["value"])
    assert r.shape == want.shape, f"{case['name']}: shape {tuple(r.shape)} != {tuple(want.shape)}"
    assert torch.equal(r, want), f"{case['name']}: got {r.flatten().tolist()} want {want.flatten().tolist()}"
    if "overflow" in exp and ex.get("overflow") is not None:
        assert ex["overflow"] == exp["overflow"], f"{case['name']}: overflow {ex['overflow']} != {exp['overflow']}"
    return "ok"
