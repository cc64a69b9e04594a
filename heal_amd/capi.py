"""ctypes binding of the HEAL MI355X C ABI (include/heal_amd.h).

Tensors are torch tensors on a ROCm device; each call passes a view descriptor
(device pointer, dtype, dims, strides) and the current torch stream.  Status
codes become the Python exceptions that pybind11 raises for the reference's
C++ exceptions: std::invalid_argument -> ValueError, std::out_of_range ->
IndexError, std::runtime_error -> RuntimeError, std::bad_alloc -> MemoryError.

The library must have been built (``python -m heal_amd.build``); importing this
module without it raises ImportError — there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import torch

HX_MAX_DIMS = 8
HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("HEAL_AMD_LIB", os.path.join(HERE, "libheal_amd.so"))

HX_INT32, HX_INT64, HX_FLOAT64 = 0, 1, 2
HX_OK, HX_ERR_INVALID_ARGUMENT, HX_ERR_OUT_OF_RANGE, HX_ERR_RUNTIME, HX_ERR_BAD_ALLOC = range(5)
HX_PERMS_TRUSTED = 1
HX_PERMS_ROWS_SAME = 2

_DTYPE = {torch.int32: HX_INT32, torch.int64: HX_INT64, torch.float64: HX_FLOAT64}


class HxView(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("dtype", ctypes.c_int32),
        ("ndim", ctypes.c_int32),
        ("dims", ctypes.c_int64 * HX_MAX_DIMS),
        ("strides", ctypes.c_int64 * HX_MAX_DIMS),
    ]


class HxDeviceStatus(ctypes.Structure):
    _fields_ = [("gdecomp_overflow", ctypes.c_uint64), ("index_out_of_range", ctypes.c_uint64)]


HX_GDECOMP_LOG_MAX = 64


class HxGdecompLog(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint64), ("tag", ctypes.c_uint64 * HX_GDECOMP_LOG_MAX),
                ("index", ctypes.c_uint64 * HX_GDECOMP_LOG_MAX)]


VP = ctypes.POINTER(HxView)
I64 = ctypes.c_int64
U64 = ctypes.c_uint64
STREAM = ctypes.c_void_p

# name -> argtypes (all return int status)
_PROTOS = {
    "hx_status_poll": [STREAM, ctypes.POINTER(HxDeviceStatus)],
    "hx_status_poll_log": [STREAM, ctypes.POINTER(HxDeviceStatus), ctypes.POINTER(HxGdecompLog)],
    "hx_malloc": [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, STREAM],
    "hx_free": [ctypes.c_void_p, STREAM],
    "hx_memset_zero": [ctypes.c_void_p, ctypes.c_size_t, STREAM],
    "hx_memcpy_h2d": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, STREAM],
    "hx_memcpy_d2h": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, STREAM],
    "hx_memcpy_d2d": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, STREAM],
    "hx_stream_sync": [STREAM],
    "hx_stream_capturing": [STREAM, ctypes.POINTER(ctypes.c_int)],
    "hx_strided_copy": [VP, VP, STREAM],
    "hx_modmul_ttt": [VP, VP, VP, VP, STREAM],
    "hx_modmul_ttc": [VP, VP, I64, VP, STREAM],
    "hx_modmul_tct": [VP, I64, VP, VP, STREAM],
    "hx_modmul_tcc": [VP, I64, I64, VP, STREAM],
    "hx_modsum_ttt": [VP, VP, VP, VP, STREAM],
    "hx_modsum_ttc": [VP, VP, I64, VP, STREAM],
    "hx_modsum_tct": [VP, I64, VP, VP, STREAM],
    "hx_modsum_tcc": [VP, I64, I64, VP, STREAM],
    "hx_mod_tt": [VP, VP, VP, STREAM],
    "hx_mod_tc": [VP, I64, VP, STREAM],
    "hx_mod_ct": [I64, VP, VP, STREAM],
    "hx_axis_modsum": [VP, VP, VP, I64, STREAM],
    "hx_modneg_tt": [VP, VP, VP, STREAM],
    "hx_modneg_tc": [VP, I64, VP, STREAM],
    "hx_modmul_axis_sum": [VP, VP, VP, VP, I64, STREAM],
    "hx_abs": [VP, VP, STREAM],
    "hx_rot_mac": [VP, VP, VP, VP, VP, I64, VP, STREAM],
    "hx_set_const": [VP, I64, ctypes.c_double, STREAM],
    "hx_g_decomposition": [VP, VP, U64, U64, STREAM],
    "hx_g_decomposition_tagged": [VP, VP, U64, U64, U64, STREAM],
    "hx_ntt": [VP, VP, VP, VP, VP, VP, VP, STREAM],
    "hx_intt": [VP, VP, VP, VP, VP, VP, VP, VP, STREAM],
    "hx_permute": [VP, VP, VP, I64, I64, STREAM],
    "hx_permute_ex": [VP, VP, VP, I64, I64, ctypes.c_uint32, STREAM],
    "hx_perms_rows_same": [VP, ctypes.POINTER(ctypes.c_int), STREAM],
    "hx_ntt_plan_create": [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, VP, VP, VP, VP, STREAM],
    "hx_ntt_plan_exec": [ctypes.c_void_p, VP, VP, STREAM],
    "hx_ntt_plan_destroy": [ctypes.c_void_p],
    "hx_ntt_plan_workspace_bytes": [ctypes.c_void_p, VP, VP, ctypes.POINTER(ctypes.c_size_t)],
    "hx_ntt_plan_exec_ws": [ctypes.c_void_p, VP, VP, ctypes.c_void_p, ctypes.c_size_t, STREAM],
    "hx_ntt_plan_exec_noalloc": [ctypes.c_void_p, VP, VP, STREAM],
    "hx_butterfly_rate": [ctypes.POINTER(ctypes.c_double), STREAM],
    "hx_keyswitch_plan_create": [ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p, ctypes.c_void_p, VP, U64, U64,
                                 STREAM],
    "hx_keyswitch_key_bytes": [ctypes.c_void_p, I64, ctypes.POINTER(ctypes.c_size_t)],
    "hx_keyswitch_plan_config": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                 ctypes.POINTER(ctypes.c_int32)],
    "hx_keyswitch_prepare_key": [ctypes.c_void_p, VP, ctypes.c_void_p, STREAM],
    "hx_keyswitch_exec": [ctypes.c_void_p, VP, ctypes.c_void_p, ctypes.c_size_t, VP, STREAM],
    "hx_keyswitch_plan_destroy": [ctypes.c_void_p],
    "hx_keyswitch_exec_marked": [ctypes.c_void_p, VP, ctypes.c_void_p, ctypes.c_size_t, VP,
                                 ctypes.POINTER(ctypes.c_void_p), STREAM],
    "hx_keyswitch_exec_ex": [ctypes.c_void_p, VP, ctypes.c_void_p, ctypes.c_size_t, I64, I64, ctypes.c_void_p, VP,
                             STREAM],
    "hx_keyswitch_row_map": [ctypes.c_void_p, VP, ctypes.c_void_p, STREAM],
    "hx_keyswitch_exec_rmac": [ctypes.c_void_p, VP, ctypes.c_void_p, ctypes.c_size_t, I64, I64, ctypes.c_void_p, VP,
                               VP, VP, I64, STREAM],
    "hx_keyswitch_group_create": [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p), VP,
                                  ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_size_t), VP, ctypes.c_int,
                                  STREAM],
    "hx_keyswitch_group_exec": [ctypes.c_void_p, STREAM],
    "hx_ntt_mfma_plan_create": [ctypes.POINTER(ctypes.c_void_p), I64, ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_void_p, ctypes.c_void_p, STREAM],
    "hx_ntt_mfma_exec": [ctypes.c_void_p, VP, VP, STREAM],
    "hx_ntt_mfma_plan_destroy": [ctypes.c_void_p],
    "hx_keyswitch_group_destroy": [ctypes.c_void_p],
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"HEAL MI355X library not found at {LIB_PATH}; build it with `python -m heal_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in _PROTOS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.hx_last_error.restype = ctypes.c_char_p
    lib.hx_last_error.argtypes = []
    lib.hx_abi_version.restype = ctypes.c_int
    return lib


lib = _load()


def exported_symbols() -> list[str]:
    return sorted(list(_PROTOS) + ["hx_last_error", "hx_abi_version"])


class HealError(RuntimeError):
    pass


def check(code: int) -> None:
    if code == HX_OK:
        return
    msg = lib.hx_last_error().decode(errors="replace")
    if code == HX_ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    if code == HX_ERR_OUT_OF_RANGE:
        raise IndexError(msg)
    if code == HX_ERR_BAD_ALLOC:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def view(t: torch.Tensor | None):
    """View descriptor of a tensor (None -> NULL pointer)."""
    if t is None:
        return None
    if t.dim() > HX_MAX_DIMS:
        raise ValueError("tensors of more than 8 dims are not supported")
    if t.dtype not in _DTYPE:
        raise TypeError(f"unsupported dtype {t.dtype}")
    v = HxView()
    v.data = t.data_ptr() if t.numel() else 0
    v.dtype = _DTYPE[t.dtype]
    v.ndim = t.dim()
    for i, (d, s) in enumerate(zip(t.shape, t.stride())):
        v.dims[i] = d
        v.strides[i] = s
    return ctypes.byref(v)


def stream_of(t: torch.Tensor | None = None) -> int:
    dev = t.device if t is not None and t.is_cuda else None
    return torch.cuda.current_stream(dev).cuda_stream


def _need_cuda(*ts):
    for t in ts:
        if isinstance(t, torch.Tensor) and not t.is_cuda:
            raise ValueError("HEAL MI355X ops take device tensors")


# ------------------------------------------------------------ op wrappers
def modmul_ttt(a, b, p, out):
    _need_cuda(a, b, p, out)
    check(lib.hx_modmul_ttt(view(a), view(b), view(p), view(out), stream_of(out)))
    return out


def modmul_ttc(a, b, p: int, out):
    _need_cuda(a, b, out)
    check(lib.hx_modmul_ttc(view(a), view(b), int(p), view(out), stream_of(out)))
    return out


def modmul_tct(a, b: int, p, out):
    _need_cuda(a, p, out)
    check(lib.hx_modmul_tct(view(a), int(b), view(p), view(out), stream_of(out)))
    return out


def modmul_tcc(a, b: int, p: int, out):
    _need_cuda(a, out)
    check(lib.hx_modmul_tcc(view(a), int(b), int(p), view(out), stream_of(out)))
    return out


def modsum_ttt(a, b, p, out):
    _need_cuda(a, b, p, out)
    check(lib.hx_modsum_ttt(view(a), view(b), view(p), view(out), stream_of(out)))
    return out


def modsum_ttc(a, b, p: int, out):
    _need_cuda(a, b, out)
    check(lib.hx_modsum_ttc(view(a), view(b), int(p), view(out), stream_of(out)))
    return out


def modsum_tct(a, b: int, p, out):
    _need_cuda(a, p, out)
    check(lib.hx_modsum_tct(view(a), int(b), view(p), view(out), stream_of(out)))
    return out


def modsum_tcc(a, b: int, p: int, out):
    _need_cuda(a, out)
    check(lib.hx_modsum_tcc(view(a), int(b), int(p), view(out), stream_of(out)))
    return out


def mod_tt(a, b, out):
    _need_cuda(a, b, out)
    check(lib.hx_mod_tt(view(a), view(b), view(out), stream_of(out)))
    return out


def mod_tc(a, b: int, out):
    _need_cuda(a, out)
    check(lib.hx_mod_tc(view(a), int(b), view(out), stream_of(out)))
    return out


def mod_ct(a: int, b, out):
    _need_cuda(b, out)
    check(lib.hx_mod_ct(int(a), view(b), view(out), stream_of(out)))
    return out


def axis_modsum(a, p, out, axis: int):
    _need_cuda(a, p, out)
    check(lib.hx_axis_modsum(view(a), view(p), view(out), int(axis), stream_of(out)))
    return out


def modneg_tt(a, p, out):
    _need_cuda(a, p, out)
    check(lib.hx_modneg_tt(view(a), view(p), view(out), stream_of(out)))
    return out


def modneg_tc(a, p: int, out):
    _need_cuda(a, out)
    check(lib.hx_modneg_tc(view(a), int(p), view(out), stream_of(out)))
    return out


def modmul_axis_sum(a, b, p, out, axis: int):
    _need_cuda(a, b, p, out)
    check(lib.hx_modmul_axis_sum(view(a), view(b), view(p), view(out), int(axis), stream_of(out)))
    return out


def rot_mac(acc, ks, ct, perm, diags, diag: int, p):
    """One diagonal rotate-multiply-accumulate step (hx_rot_mac); ks=None for diagonal 0."""
    _need_cuda(acc, ct, diags, p)
    check(lib.hx_rot_mac(view(acc), view(ks), view(ct), view(perm), view(diags), int(diag), view(p), stream_of(acc)))
    return acc


def abs_(a, out):
    _need_cuda(a, out)
    check(lib.hx_abs(view(a), view(out), stream_of(out)))
    return out


def set_const(a, value):
    _need_cuda(a)
    if a.dtype == torch.float64:
        check(lib.hx_set_const(view(a), 0, float(value), stream_of(a)))
    else:
        check(lib.hx_set_const(view(a), int(value), 0.0, stream_of(a)))
    return a


def g_decomposition(a, out, power: int, base_bits: int, tag: int = 0):
    """tag (optional): carried by the elements of this call in the overflow log (status_poll)."""
    _need_cuda(a, out)
    if tag:
        check(lib.hx_g_decomposition_tagged(view(a), view(out), int(power), int(base_bits), int(tag),
                                            stream_of(out)))
    else:
        check(lib.hx_g_decomposition(view(a), view(out), int(power), int(base_bits), stream_of(out)))
    return out


def ntt(a, p, perm, twiddles, out, log2p=None, mu=None):
    _need_cuda(a, p, perm, twiddles, out)
    check(lib.hx_ntt(view(a), view(p), view(perm), view(twiddles), view(log2p), view(mu), view(out), stream_of(out)))
    return out


def intt(a, p, perm, inv_twiddles, m_inv, out, log2p=None, mu=None):
    _need_cuda(a, p, perm, inv_twiddles, m_inv, out)
    check(lib.hx_intt(view(a), view(p), view(perm), view(inv_twiddles), view(m_inv), view(log2p), view(mu),
                      view(out), stream_of(out)))
    return out


def permute(a, perms, out, elementwise_axis: int, perm_axis: int, trusted: bool = False, rows_same: bool = False):
    """trusted: perms was validated by an earlier call.  rows_same (with trusted):
    every row of perms equals row 0, as perms_rows_same reports."""
    _need_cuda(a, perms, out)
    if trusted:
        flags = HX_PERMS_TRUSTED | (HX_PERMS_ROWS_SAME if rows_same else 0)
        check(lib.hx_permute_ex(view(a), view(perms), view(out), int(elementwise_axis), int(perm_axis),
                                flags, stream_of(out)))
    else:
        check(lib.hx_permute(view(a), view(perms), view(out), int(elementwise_axis), int(perm_axis), stream_of(out)))
    return out


def perms_rows_same(perms) -> bool:
    """True when every row of the 2-D table perms equals its row 0 (synchronous)."""
    _need_cuda(perms)
    same = ctypes.c_int(0)
    check(lib.hx_perms_rows_same(view(perms), ctypes.byref(same), stream_of(perms)))
    return bool(same.value)


def strided_copy(src, dst):
    _need_cuda(src, dst)
    check(lib.hx_strided_copy(view(src), view(dst), stream_of(dst)))
    return dst


def stream_capturing(stream: int | None = None) -> bool:
    c = ctypes.c_int(0)
    check(lib.hx_stream_capturing(stream if stream is not None else stream_of(), ctypes.byref(c)))
    return bool(c.value)


def butterfly_rate(stream: int | None = None) -> float:
    """Butterflies per second of the NTT passes' 60-bit lazy butterfly with its
    operands in registers (the arithmetic ceiling of a pass on this GPU)."""
    out = ctypes.c_double(0)
    check(lib.hx_butterfly_rate(ctypes.byref(out), stream if stream is not None else stream_of()))
    return out.value


def status_poll(stream: int | None = None) -> dict:
    """Takes the device counters and the g_decomposition overflow log:
    gdecomp_log lists (tag, row-major element index) pairs, sorted."""
    st = HxDeviceStatus()
    log = HxGdecompLog()
    check(lib.hx_status_poll_log(stream if stream is not None else stream_of(), ctypes.byref(st), ctypes.byref(log)))
    entries = sorted((int(log.tag[i]), int(log.index[i])) for i in range(log.count))
    return {"gdecomp_overflow": st.gdecomp_overflow, "index_out_of_range": st.index_out_of_range,
            "gdecomp_log": entries}


class NttPlan:
    """Device snapshot of (p, perm, twiddles[, m_inv]) for repeated transforms."""

    def __init__(self, p, perm, twiddles, m_inv=None, inverse: bool = False):
        _need_cuda(p, perm, twiddles, m_inv)
        self._h = ctypes.c_void_p()
        self.inverse = bool(inverse)
        check(lib.hx_ntt_plan_create(ctypes.byref(self._h), int(self.inverse), view(p), view(perm), view(twiddles),
                                     view(m_inv), stream_of(p)))

    def __call__(self, a, out):
        _need_cuda(a, out)
        check(lib.hx_ntt_plan_exec(self._h, view(a), view(out), stream_of(out)))
        return out

    def workspace_bytes(self, a, out) -> int:
        """Workspace that exec_ws needs for these views."""
        n = ctypes.c_size_t()
        check(lib.hx_ntt_plan_workspace_bytes(self._h, view(a), view(out), ctypes.byref(n)))
        return n.value

    def exec_ws(self, a, out, ws):
        """The transform with a caller's workspace (a device tensor of at least
        workspace_bytes(a, out) bytes, 256-byte aligned)."""
        _need_cuda(a, out, ws)
        check(lib.hx_ntt_plan_exec_ws(self._h, view(a), view(out), ws.data_ptr(), ws.numel() * ws.element_size(),
                                      stream_of(out)))
        return out

    def exec_noalloc(self, a, out):
        """The transform with the library's fixed scratch (no device allocation;
        calls overlapping on one GPU race)."""
        _need_cuda(a, out)
        check(lib.hx_ntt_plan_exec_noalloc(self._h, view(a), view(out), stream_of(out)))
        return out

    def close(self):
        if self._h:
            check(lib.hx_ntt_plan_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class KeySwitchPlan:
    """Fused hybrid key-switch over a forward and an inverse NttPlan of the same
    moduli and permutation (hx_keyswitch_* in include/heal_amd.h)."""

    def __init__(self, fwd: NttPlan, inv: NttPlan, perms, dnum: int, base_bits: int):
        _need_cuda(perms)
        self._h = ctypes.c_void_p()
        self._plans = (fwd, inv)  # the fused plan borrows them
        self.dnum, self.base_bits = int(dnum), int(base_bits)
        check(lib.hx_keyswitch_plan_create(ctypes.byref(self._h), fwd._h, inv._h, view(perms), self.dnum,
                                           self.base_bits, stream_of(perms)))

    def config(self) -> dict:
        """The tuned shape: index bits of the middle kernel, its limb group, and
        the columns per workgroup of the first and last kernels."""
        lo, g, t = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(lib.hx_keyswitch_plan_config(self._h, ctypes.byref(lo), ctypes.byref(g), ctypes.byref(t)))
        return {"low_bits": lo.value, "limb_group": g.value, "tile": t.value}

    def prepare_key(self, evk):
        """evk [dnum, B, 2, m, k] -> device-layout key (an int64 tensor)."""
        _need_cuda(evk)
        nbytes = ctypes.c_size_t()
        check(lib.hx_keyswitch_key_bytes(self._h, int(evk.shape[1]), ctypes.byref(nbytes)))
        key = torch.empty(nbytes.value // 8, dtype=torch.int64, device=evk.device)
        check(lib.hx_keyswitch_prepare_key(self._h, view(evk), key.data_ptr(), stream_of(evk)))
        return key

    def __call__(self, x, key, out):
        _need_cuda(x, key, out)
        check(lib.hx_keyswitch_exec(self._h, view(x), key.data_ptr(), key.numel() * key.element_size(), view(out),
                                    stream_of(out)))
        return out

    def exec_marked(self, x, key, out, events):
        """The key-switch with torch.cuda.Events recorded before the first kernel
        and after each of the three (events: four recorded-once Events or None)."""
        _need_cuda(x, key, out)
        ev = (ctypes.c_void_p * 4)(*[e.cuda_event if e is not None else None for e in events])
        check(lib.hx_keyswitch_exec_marked(self._h, view(x), key.data_ptr(), key.numel() * key.element_size(),
                                           view(out), ev, stream_of(out)))
        return out

    def exec_ex(self, x, key, out, key_limbs: int, key_count: int, row_map=None):
        """Key-switch with a shared or per-instance key of key_limbs limbs and an
        optional row map (int32 [m], from row_map()) for the Galois permutation."""
        _need_cuda(x, key, out)
        check(lib.hx_keyswitch_exec_ex(self._h, view(x), key.data_ptr(), key.numel() * key.element_size(),
                                       int(key_limbs), int(key_count),
                                       row_map.data_ptr() if row_map is not None else None, view(out),
                                       stream_of(out)))
        return out

    def exec_rmac(self, x, key, acc, ct, diags, diag: int, key_limbs: int, key_count: int, row_map=None):
        """exec_ex of x fused with one rotate-multiply-accumulate step into acc
        (hx_keyswitch_exec_rmac): acc_0 += pt * (KS_0 + ct_0 permuted), acc_1 +=
        pt * KS_1, pt the diagonal `diag` of diags [B, M, M, k]."""
        _need_cuda(x, key, acc, ct, diags)
        check(lib.hx_keyswitch_exec_rmac(self._h, view(x), key.data_ptr(), key.numel() * key.element_size(),
                                         int(key_limbs), int(key_count),
                                         row_map.data_ptr() if row_map is not None else None, view(acc), view(ct),
                                         view(diags), int(diag), stream_of(acc)))
        return acc

    def row_map(self, perm):
        """int32 [m] row map of one Galois permutation perm [m]."""
        _need_cuda(perm)
        rm = torch.empty(perm.shape[-1], dtype=torch.int32, device=perm.device)
        check(lib.hx_keyswitch_row_map(self._h, view(perm), rm.data_ptr(), stream_of(perm)))
        return rm

    def close(self):
        if self._h:
            check(lib.hx_keyswitch_plan_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _view_struct(t):
    v = HxView()
    v.data = t.data_ptr() if t.numel() else 0
    v.dtype = _DTYPE[t.dtype]
    v.ndim = t.dim()
    for i, (d, st) in enumerate(zip(t.shape, t.stride())):
        v.dims[i] = d
        v.strides[i] = st
    return v


class KeySwitchGroup:
    """Key-switches of mixed (m, k) run together (hx_keyswitch_group_*): member i
    is plans[i] on xs[i] with keys[i] into outs[i], fixed at creation."""

    def __init__(self, plans, xs, keys, outs):
        n = len(plans)
        if not (len(xs) == len(keys) == len(outs) == n):
            raise ValueError("plans, xs, keys and outs differ in length")
        _need_cuda(*xs, *keys, *outs)
        self._keep = (list(plans), list(xs), list(keys), list(outs))
        ph = (ctypes.c_void_p * n)(*[p._h.value for p in plans])
        xv = (HxView * n)(*[_view_struct(x) for x in xs])
        ov = (HxView * n)(*[_view_struct(o) for o in outs])
        kp = (ctypes.c_void_p * n)(*[k.data_ptr() for k in keys])
        kb = (ctypes.c_size_t * n)(*[k.numel() * k.element_size() for k in keys])
        self._h = ctypes.c_void_p()
        check(lib.hx_keyswitch_group_create(ctypes.byref(self._h), ph, xv, kp, kb, ov, n, stream_of(outs[0])))

    def __call__(self):
        check(lib.hx_keyswitch_group_exec(self._h, stream_of()))

    def close(self):
        if self._h:
            check(lib.hx_keyswitch_group_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
