"""Loaders for the parity checkers under oracle/ (test infrastructure).

`oracle()` returns the C restatement (oracle/liboracle.so, built by
`make -C oracle`); `reference()` returns the compiled reference harness
(oracle/_ref/libheal_ref.so) or None where it is not built.  Both take
hx_view descriptors over HOST memory; `hview()` makes one from a CPU torch
tensor.  Status codes follow include/heal_amd.h.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")
REF_LIB = os.path.join(ORACLE_DIR, "_ref", "libheal_ref.so")

HX_MAX_DIMS = 8
CODES = {0: None, 1: "invalid_argument", 2: "out_of_range", 3: "runtime_error", 4: "bad_alloc"}
_DT = {torch.int32: 0, torch.int64: 1, torch.float64: 2}


class HView(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("dtype", ctypes.c_int32),
        ("ndim", ctypes.c_int32),
        ("dims", ctypes.c_int64 * HX_MAX_DIMS),
        ("strides", ctypes.c_int64 * HX_MAX_DIMS),
    ]


def hview(t):
    if t is None:
        return None
    assert not t.is_cuda
    v = HView()
    v.data = t.data_ptr() if t.numel() else 0
    v.dtype = _DT[t.dtype]
    v.ndim = t.dim()
    for i, (d, s) in enumerate(zip(t.shape, t.stride())):
        v.dims[i] = d
        v.strides[i] = s
    return ctypes.byref(v)


VP = ctypes.POINTER(HView)
I64, U64 = ctypes.c_int64, ctypes.c_uint64

_ORACLE_PROTOS = {
    "or_modop": [ctypes.c_int, VP, VP, I64, VP, I64, VP],
    "or_mod": [VP, I64, VP, I64, VP, ctypes.c_int, ctypes.c_int],
    "or_axis_modsum": [VP, VP, VP, I64],
    "or_g_decomposition": [VP, VP, U64, U64, ctypes.POINTER(U64)],
    "or_ntt": [VP, VP, VP, VP, VP],
    "or_intt": [VP, VP, VP, VP, VP, VP, ctypes.c_int],
    "or_permute": [VP, VP, VP, I64, I64],
    "or_strided_copy": [VP, VP],
}

_REF_PROTOS = {
    "ref_modop": [ctypes.c_int, VP, VP, I64, VP, I64, VP],
    "ref_mod": [ctypes.c_int, VP, I64, VP, I64, VP],
    "ref_axis_modsum": [VP, VP, VP, I64],
    "ref_g_decomposition": [VP, VP, U64, U64],
    "ref_ntt": [VP, VP, VP, VP, VP],
    "ref_intt": [VP, VP, VP, VP, VP, VP],
    "ref_permute": [VP, VP, VP, I64, I64],
    "ref_make_contiguous": [VP, VP],
    "ref_keyswitch": [VP, VP, VP, VP, VP, VP, VP, VP, U64, U64, VP, ctypes.POINTER(ctypes.c_double)],
}

_cache = {}


def _bind(lib, protos):
    for name, args in protos.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    return lib


def oracle():
    if "oracle" not in _cache:
        src = os.path.join(ORACLE_DIR, "heal_oracle.c")
        if not os.path.exists(ORACLE_LIB) or os.path.getmtime(ORACLE_LIB) < os.path.getmtime(src):
            subprocess.run(["make", "-s", "-C", ORACLE_DIR, "liboracle.so"], check=True)
        lib = _bind(ctypes.CDLL(ORACLE_LIB), _ORACLE_PROTOS)
        lib.or_last_error.restype = ctypes.c_char_p
        _cache["oracle"] = lib
    return _cache["oracle"]


def reference():
    if "ref" not in _cache:
        lib = None
        if os.path.exists(REF_LIB):
            lib = _bind(ctypes.CDLL(REF_LIB), _REF_PROTOS)
            lib.ref_last_error.restype = ctypes.c_char_p
        _cache["ref"] = lib
    return _cache["ref"]
