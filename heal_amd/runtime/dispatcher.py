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
call, since no transcript in the reference fixes another.
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
        """(a * b) % q summed over `axis` (a modmul into a temporary, then axis_modsum)."""
        prod = self._f("allocate_on_hardware", a)(self._broadcast_shape(a.shape, b.shape))
        self._f("modmul_ttt", a)(a, b, q_list, prod)
        self._f("axis_modsum", a)(prod, q_list, out, axis)
        return out

    @staticmethod
    def _broadcast_shape(x, y):
        n = max(len(x), len(y))
        x = [1] * (n - len(x)) + list(x)
        y = [1] * (n - len(y)) + list(y)
        return [max(i, j) for i, j in zip(x, y)]

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
        # ops the transcript format names whose semantics no part of the
        # reference defines (moveaxis, pad_single_axis, get_slice, abs,
        # set_const_val, _modneg_*)
        raise NotImplementedError(f"transcript op {name!r} is not part of the HEAL hardware API")
