"""RNS-NTT parameters in the layout the HEAL API consumes.

For a ring degree N (the NTT length m of include/ntt.h in the reference) and
primes q_t = 1 mod 2N this builds, per limb t:
  - psi_t, a primitive 2N-th root of unity mod q_t;
  - twiddles[t, i] = psi_t^bitrev(i) and inv_twiddles[t, i] = psi_t^-bitrev(i),
    the tables the reference's butterfly networks index as tw[t, stage + u]
    and itw[t, half + group] (example_impl/ntt_impl.cpp:99, 153);
  - m_inv[t] = N^-1 mod q_t;
and perm = bit reversal of log2(N) bits.  With these tables the forward
transform is out[u] = sum_i a_i psi^((2u+1) i) mod q (negacyclic, natural
order).  Galois maps for permute() are built here as well.
"""
from __future__ import annotations

import random

import numpy as np
import torch

_MR_BASES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)


def is_prime(n: int) -> bool:
    """Deterministic Miller–Rabin for n < 3.3e24 (fixed bases 2..37)."""
    if n < 2:
        return False
    for p in _MR_BASES:
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in _MR_BASES:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def primes_below(bits: int, modulus: int, count: int) -> list[int]:
    """The `count` largest primes q < 2^bits with q = 1 mod `modulus`, descending."""
    out = []
    q = ((1 << bits) - 1) // modulus * modulus + 1
    while len(out) < count:
        if q < 2:
            raise ValueError("not enough primes")
        if q < (1 << bits) and is_prime(q):
            out.append(q)
        q -= modulus
    return out


def ntt_primes(n: int, bit_sizes: list[int]) -> list[int]:
    """Distinct NTT-friendly primes (q = 1 mod 2n), one per requested bit size.
    Repeated sizes take the next smaller prime of that size."""
    used: dict[int, int] = {}
    out = []
    for b in bit_sizes:
        k = used.get(b, 0)
        out.append(primes_below(b, 2 * n, k + 1)[k])
        used[b] = k + 1
    return out


def keyswitch_primes(L: int = 24, n: int = 65536) -> list[int]:
    """Limb t has 60 - (t mod 11) bits (50..60), all q = 1 mod 2^17 (so also
    1 mod 2n for every n <= 2^16)."""
    return ntt_primes(max(n, 65536), [60 - (t % 11) for t in range(L)])


def bitrev(n_bits: int) -> np.ndarray:
    m = 1 << n_bits
    r = np.zeros(m, dtype=np.int64)
    idx = np.arange(m, dtype=np.int64)
    for b in range(n_bits):
        r |= ((idx >> b) & 1) << (n_bits - 1 - b)
    return r


def _factor_small(n: int) -> list[int]:
    f, d = [], 2
    while d * d <= n:
        while n % d == 0:
            if d not in f:
                f.append(d)
            n //= d
        d += 1
    if n > 1:
        f.append(n)
    return f


def primitive_root_2n(q: int, n: int, seed: int = 0) -> int:
    """A primitive 2n-th root of unity mod prime q (q = 1 mod 2n, n a power of 2)."""
    if (q - 1) % (2 * n):
        raise ValueError("q != 1 mod 2n")
    rng = random.Random(seed * 1000003 + q)
    e = (q - 1) // (2 * n)
    while True:
        x = rng.randrange(2, q - 1)
        psi = pow(x, e, q)
        if pow(psi, n, q) == q - 1:  # order exactly 2n since n is a power of 2
            return psi


def _powers(base: int, count: int, q: int) -> list[int]:
    out = [1] * count
    for i in range(1, count):
        out[i] = out[i - 1] * base % q
    return out


class NttParams:
    """Tables for one (N, primes) pair, as int64 torch tensors on `device`."""

    def __init__(self, n: int, primes: list[int], device="cpu", dtype=torch.int64, psi: list[int] | None = None):
        if n & (n - 1):
            raise ValueError("n must be a power of 2")
        self.n = n
        self.primes = list(primes)
        self.log_n = n.bit_length() - 1
        self.psi = psi if psi is not None else [primitive_root_2n(q, n) for q in primes]
        br = bitrev(self.log_n)
        tw = np.zeros((len(primes), n), dtype=np.int64)
        itw = np.zeros((len(primes), n), dtype=np.int64)
        for t, (q, g) in enumerate(zip(primes, self.psi)):
            pw = np.array(_powers(g, n, q), dtype=object)
            ipw = np.array(_powers(pow(g, -1, q), n, q), dtype=object)
            tw[t] = pw[br].astype(np.int64)
            itw[t] = ipw[br].astype(np.int64)
        self.p = torch.tensor(self.primes, dtype=dtype, device=device)
        self.perm = torch.tensor(br, dtype=dtype, device=device)
        self.twiddles = torch.tensor(tw, dtype=dtype, device=device)
        self.inv_twiddles = torch.tensor(itw, dtype=dtype, device=device)
        self.m_inv = torch.tensor([pow(n, -1, q) for q in primes], dtype=dtype, device=device)

    def to(self, device):
        for name in ("p", "perm", "twiddles", "inv_twiddles", "m_inv"):
            setattr(self, name, getattr(self, name).to(device))
        return self


def galois_perms(n: int, galois_elt: int, limbs: int, domain: str = "coeff", device="cpu", dtype=torch.int64):
    """perms[t, j] for permute() (elementwise axis = limbs, perm axis = N).

    domain="coeff": the index part of x -> x^g on coefficients,
      out[j] = a[(j * g^-1 mod 2n) mod n] (the sign flip of the true
      automorphism is not part of a gather);
    domain="ntt": the automorphism on values in the natural-order negacyclic
      NTT domain, out[u] = a[((2u+1) g - 1)/2 mod n], which is sign-free.
    """
    two_n = 2 * n
    g = galois_elt % two_n
    j = np.arange(n, dtype=np.int64)
    if domain == "coeff":
        ginv = pow(g, -1, two_n)
        src = (j * ginv % two_n) % n
    elif domain == "ntt":
        src = (((2 * j + 1) * g % two_n) - 1) // 2 % n
    else:
        raise ValueError(domain)
    return torch.tensor(np.broadcast_to(src, (limbs, n)).copy(), dtype=dtype, device=device)


def uniform_residues(shape_prefix, primes, n, generator: torch.Generator, device="cpu"):
    """Tensor [*shape_prefix, n, L] of residues uniform in [0, q_t) (limb axis last)."""
    cols = []
    for q in primes:
        cols.append(torch.randint(0, q, (*shape_prefix, n, 1), generator=generator, dtype=torch.int64))
    return torch.cat(cols, dim=-1).to(device)
