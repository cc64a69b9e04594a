"""HEAL on MI355X: a gfx950-native backend for the HEAL hardware API.

Subpackages / modules:
  capi     ctypes binding of the C ABI (include/heal_amd.h) over torch device tensors
  params   NTT parameter generation (primes, roots, twiddle tables, Galois maps)
  keyswitch  the hybrid key-switch chain composed from HEAL calls (plain and fused forms)
  build    in-tree build of libheal_amd.so
"""
import torch  # noqa: F401  (loads the ROCm runtime before the HIP library)

__all__ = ["capi", "params", "keyswitch", "build"]
