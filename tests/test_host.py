"""CPU tests of host-side product code: the modular arithmetic core compiled
for the host, NTT parameter generation, and the C ABI library's exports."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import pytest
import torch

from heal_amd import params as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_arith_core_matches_int128(tmp_path):
    """heal_amd/csrc/hx_arith.hpp on the host against the compiler's exact 128-bit %."""
    exe = tmp_path / "arith_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe), os.path.join(ROOT, "tests", "arith_host_check.cpp")],
                   check=True)
    out = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True)
    assert out.returncode == 0, out.stdout
    assert out.stdout.startswith("OK")


@pytest.mark.parametrize("n,bits", [(4096, [30] * 4), (65536, [60, 59, 55, 50])])
def test_ntt_params_consistent(n, bits):
    primes = P.ntt_primes(n, bits)
    assert len(set(primes)) == len(primes)
    prm = P.NttParams(n, primes)
    br = P.bitrev(prm.log_n)
    assert sorted(br.tolist()) == list(range(n))
    for t, (q, psi) in enumerate(zip(primes, prm.psi)):
        assert P.is_prime(q) and q % (2 * n) == 1 and q < 2 ** bits[t]
        assert pow(psi, n, q) == q - 1
        tw = prm.twiddles[t].tolist()
        itw = prm.inv_twiddles[t].tolist()
        for i in (0, 1, 2, n // 2, n - 1):
            assert tw[i] == pow(psi, int(br[i]), q)
            assert tw[i] * itw[i] % q == 1
        assert prm.m_inv[t].item() * n % q == 1


def test_keyswitch_primes_distinct_and_sized():
    ps = P.keyswitch_primes(24)
    assert len(set(ps)) == 24
    for t, q in enumerate(ps):
        assert q.bit_length() == 60 - (t % 11) and q % (1 << 17) == 1 and P.is_prime(q)


def test_galois_perms_are_permutations():
    for dom in ("coeff", "ntt"):
        pm = P.galois_perms(1024, 5, 3, domain=dom)
        assert pm.shape == (3, 1024)
        assert sorted(pm[0].tolist()) == list(range(1024))


def test_galois_ntt_domain_is_the_automorphism():
    """In the NTT domain, permuting by the 'ntt' map equals transforming a(x^g)."""
    n, g = 16, 5
    q = P.ntt_primes(n, [30])[0]
    prm = P.NttParams(n, [q])
    psi = prm.psi[0]
    a = list(range(3, 3 + n))
    # a(x^g) mod x^n + 1
    b = [0] * n
    for i, c in enumerate(a):
        e = i * g % (2 * n)
        if e >= n:
            b[e - n] = (b[e - n] - c) % q
        else:
            b[e] = (b[e] + c) % q
    ev = lambda poly: [sum(poly[i] * pow(psi, (2 * u + 1) * i, q) for i in range(n)) % q for u in range(n)]
    A, B = ev(a), ev(b)
    perm = P.galois_perms(n, g, 1, domain="ntt")[0].tolist()
    assert [A[perm[u]] for u in range(n)] == B


def _declared_symbols():
    with open(os.path.join(ROOT, "include", "heal_amd.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hx_[a-z0-9_]+)\s*\(", text)))


def test_capi_library_exports_every_declared_symbol():
    lib_path = os.path.join(ROOT, "heal_amd", "libheal_amd.so")
    if not os.path.exists(lib_path):
        from heal_amd import build

        build.build_all()
    lib = ctypes.CDLL(lib_path)
    syms = _declared_symbols()
    assert "hx_ntt" in syms and "hx_permute" in syms
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    lib.hx_abi_version.restype = ctypes.c_int
    assert lib.hx_abi_version() == 1


def test_capi_errors_without_gpu_work():
    """Argument validation happens before any device call."""
    from heal_amd import capi

    with pytest.raises(ValueError):
        capi.check(capi.lib.hx_modmul_ttt(None, None, None, None, None))
    with pytest.raises(ValueError):
        capi.check(capi.lib.hx_mod_tt(None, None, None, None))


def test_bench_kernel_fingerprint_tracks_code_not_comments():
    """bench.py's profiled traffic and instruction counts carry the fingerprint
    of the kernel sources they were measured on: comment and layout edits keep
    it, a code change does not (the bench then reports the constants as stale)."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    code = "int f(int x) {\n    return x + 1;  // one more\n}\n"
    same = "/* doc */ int f(int x)\n{ return x + 1; }\n// trailing note\n"
    other = "int f(int x) {\n    return x + 2;\n}\n"
    assert bench.source_code_only(code) == bench.source_code_only(same)
    assert bench.source_code_only(code) != bench.source_code_only(other)
    fp = bench.kernel_fingerprint()
    assert len(fp) == 16 and all(len(e["kernels"]) == 16 for e in bench.PROFILED.values())
