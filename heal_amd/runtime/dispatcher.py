"""Op-name dispatch of transcript instructions onto the lattica_hw module.

Method names and argument orders of the ops the reference's dispatcher
implements are kept (python_execution/lattica_heal_runtime/py_wrapper.py:
114-181): host_to_device(tensor, dtype), device_to_host(a), empty(shape,
dtype), axis_modsum(a, axis, q_list, out), reshape(t, shape),
_modmul_*/_modsum_*(a, b, p, out), expand(a, repeat, axis), squeeze,
unsqueeze, contiguous, and ntt(a, perm, perm_pairs, q_list, log2p, mu_list,
psi_arr, out, tile, skip_perm).  The reference declares but does not implement
intt, take_along_axis, apply_g_decomp and _mod_* (device_interface.py:44-76);
here they are implemented with the argument order of the corresponding C++ API
call, since no transcript in the reference fixes another.  The same holds for
the ops it names without any definition (_modneg_*, modmul_axis_sum, abs,
set_const_val, moveaxis, pad_single_axis, get_slice): their argument orders
and semantics are this library's (lattica_hw_api.h).
"""
from __future__ import annotations

import torch

_SFX = {torch.int32: "32", torch.int64: "64", torch.float64: "float64"}


class HealDispatcher:
    def __init__(self, module=None):
        if module is None:
            from heal_amd import lattica_hw as module
        self.m = module
        self._cls = {getattr(module, f"DeviceTensor{s}"): s for s in ("32", "64", "float64")}

    # -- helpers --------------------------------------------------------------
    def _sfx(self, a):
        try:
            return self._cls[type(a)]
        except KeyError:
            raise TypeError(f"Unsupported tensor type: {type(a)}") from None

    def _f(self, name, a):
        return getattr(self.m, f"{name}_{self._sfx(a)}")

    # -- memory ------------------------------------------------------------------
    def host_to_device(self, tensor, dtype):
        t = tensor if isinstance(tensor, torch.Tensor) else torch.as_tensor(tensor)
        t = t.to(dtype)
        return getattr(self.m, f"host_to_device_{_SFX[dtype]}")(t)

    def device_to_host(self, a):
        return self._f("device_to_host", a)(a)

    def empty(self, shape, dtype):
        if dtype not in (torch.int32, torch.int64):
            raise TypeError(f"Unsupported tensor type: {dtype}")
        return getattr(self.m, f"allocate_on_hardware_{_SFX[dtype]}")(list(shape))

    def zeros(self, shape, dtype):
        return getattr(self.m, f"allocate_on_hardware_{_SFX[dtype]}")(list(shape))

    # -- views -----------------------------------------------------------------------
    def reshape(self, device_tensor, new_shape):
        device_tensor.reshape(list(new_shape))
        return device_tensor

    def expand(self, a, repeat, axis):
        return self._f("expand", a)(a, axis, repeat)

    def squeeze(self, a, axis):
        return self._f("squeeze", a)(a, axis)

    def unsqueeze(self, a, axis):
        return self._f("unsqueeze", a)(a, axis)

    def contiguous(self, a):
        return self._f("make_contiguous", a)(a)

    def new_reference(self, a):
        return a

    def flatten(self, a):
        n = 1
        for d in a.shape:
            n *= d
        a.reshape([n])
        return a

    # -- modular arithmetic ------------------------------------------------------
    def _modop(self, name, a, b, p, out):
        self._f(name, a)(a, b, p, out)
        return out

    def _modmul_ttt(self, a, b, p, out):
        return self._modop("modmul_ttt", a, b, p, out)

    def _modmul_ttc(self, a, b, p, out):
        return self._modop("modmul_ttc", a, b, p, out)

    def _modmul_tct(self, a, b, p, out):
        return self._modop("modmul_tct", a, b, p, out)

    def _modmul_tcc(self, a, b, p, out):
        return self._modop("modmul_tcc", a, b, p, out)

    def _modsum_ttt(self, a, b, p, out):
        return self._modop("modsum_ttt", a, b, p, out)

    def _modsum_ttc(self, a, b, p, out):
        return self._modop("modsum_ttc", a, b, p, out)

    def _modsum_tct(self, a, b, p, out):
        return self._modop("modsum_tct", a, b, p, out)

    def _modsum_tcc(self, a, b, p, out):
        return self._modop("modsum_tcc", a, b, p, out)

    def _mod_tt(self, a, b, out):
        self._f("mod_tt", out)(a, b, out)
        return out

    def _mod_tc(self, a, b, out):
        self._f("mod_tc", out)(a, b, out)
        return out

    def _mod_ct(self, a, b, out):
        self._f("mod_ct", out)(a, b, out)
        return out

    def axis_modsum(self, a, axis, q_list, out):
        self._f("axis_modsum", a)(a, q_list, out, axis)
        return out

    def modmul_axis_sum(self, a, b, q_list, out, axis):
        """axis_modsum(modmul_ttt(a, b, q), q, axis) in one kernel (no product tensor)."""
        self._f("modmul_axis_sum", a)(a, b, q_list, out, axis)
        return out

    # -- transcript ops the reference names without defining (device_interface.py:
    #    47-48, 75-107); semantics as in lattica_hw_api.h
    def _modneg_tt(self, a, p, out):
        self._f("modneg_tt", out)(a, p, out)
        return out

    def _modneg_tc(self, a, p, out):
        self._f("modneg_tc", out)(a, p, out)
        return out

    def abs(self, a, out):
        self._f("abs", out)(a, out)
        return out

    def set_const_val(self, a, value):
        self._f("set_const_val", a)(a, value)
        return a

    def moveaxis(self, a, source, destination):
        return self._f("moveaxis", a)(a, source, destination)

    def pad_single_axis(self, a, axis, before, after):
        return self._f("pad_single_axis", a)(a, axis, before, after)

    def get_slice(self, a, key):
        """Basic indexing a[key] (ints, slices with positive step, one Ellipsis): a view."""
        shape = list(a.shape)
        key = list(key) if isinstance(key, (tuple, list)) else [key]
        if sum(k is Ellipsis for k in key) > 1:
            raise IndexError("an index can only have a single ellipsis")
        if Ellipsis in key:
            i = key.index(Ellipsis)
            key = key[:i] + [slice(None)] * (len(shape) - (len(key) - 1)) + key[i + 1:]
        if len(key) > len(shape):
            raise IndexError("too many indices for tensor")
        spec = []
        for k, dim in zip(key, shape):
            if isinstance(k, slice):
                start, stop, step = k.indices(dim)
                if step <= 0:
                    raise ValueError("slice step must be positive")
                stop = max(stop, start)
                spec.append((start, stop, step, False))
            else:
                i = int(k)
                if i < -dim or i >= dim:
                    raise IndexError(f"index {i} is out of bounds for dimension with size {dim}")
                spec.append((i + dim if i < 0 else i, 0, 1, True))
        return self._f("get_slice", a)(a, spec)

    # -- special ops ---------------------------------------------------------------
    def ntt(self, a, perm, perm_pairs, q_list, log2p, mu_list, psi_arr, out, tile, skip_perm):
        if skip_perm:
            raise NotImplementedError(f"skip_perm is not supported. {skip_perm=}")
        if tile:
            a = self.expand(a, 2, -1)
        self._f("ntt", a)(a, q_list, perm, psi_arr, log2p, mu_list, out)
        return out

    def intt(self, a, q_list, perm, inv_psi_arr, m_inv, log2p, mu_list, out):
        self._f("intt", a)(a, q_list, perm, inv_psi_arr, m_inv, log2p, mu_list, out)
        return out

    def take_along_axis(self, a, perms, out, elementwise_axis, perm_axis):
        self._f("permute", a)(a, perms, out, elementwise_axis, perm_axis)
        return out

    permute = take_along_axis

    def apply_g_decomp(self, a, out, power, base_bits):
        self._f("g_decomposition", a)(a, out, power, base_bits)
        return out

    def __getattr__(self, name):
        raise NotImplementedError(f"transcript op {name!r} is not part of the HEAL hardware API")
