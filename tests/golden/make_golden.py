"""Writes the golden fixtures used by the parity tests.

reference_kat.json restates, as data, the inputs and expected outputs of the
reference's unit tests (Lattica-ai/heal tests/test_*.cpp).  Literal tables are
copied as numbers; where a reference test draws unseeded random inputs and
compares with a torch formula, the fixture fixes seeded inputs and stores the
value of that same formula.  Each case cites the test it restates.

ntt_closed_form.json holds forward-NTT vectors computed independently of any
butterfly network, by the defining sum out[u] = sum_i a_i psi^((2u+1) i) mod q
(with the bit-reversed psi-power twiddle tables of heal_amd.params), at m = 16
and 64 for 30-bit and 60-bit primes.

ntt_signed_kat.json holds length-2 NTT/INTT vectors worked by hand from the
reference's loops on inputs outside [0, q), where its truncating % keeps
negative representatives (see ntt_signed_cases).

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import random

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def T(values, dtype="int32", view=None):
    a = np.array(values)
    d = {"dtype": dtype, "shape": list(a.shape), "data": a.reshape(-1).tolist()}
    if view:
        d["view"] = view
    return d


def rnd(rng, lo, hi, shape):
    return np.array([rng.randrange(lo, hi) for _ in range(int(np.prod(shape)))]).reshape(shape)


def materialize(t):
    a = np.array(t["data"], dtype=np.int64).reshape(t["shape"])
    for op in t.get("view", []):
        if op[0] == "transpose":
            a = np.swapaxes(a, op[1], op[2])
        elif op[0] == "permute":
            a = np.transpose(a, op[1])
    return a


def reference_cases():
    rng = random.Random(20240901)
    cases = []

    def add(name, cite, op, args, expect):
        cases.append({"name": name, "cite": cite, "op": op, "args": args, "expect": expect})

    def val(a, dtype="int32"):
        return {"value": T(np.asarray(a), dtype)}

    err_inv = {"error": "invalid_argument"}
    err_oor = {"error": "out_of_range"}

    # ---------------- NTT (tests/test_ntt.cpp:7-53)
    a = [[[[1, 2]], [[3, 4]], [[5, 6]], [[7, 8]]]]
    ntt_args = {
        "a": T(a), "p": T([17, 257]), "perm": T([0, 2, 1, 3]),
        "twiddles": T([[1, 4, 2, 8], [1, 16, 4, 64]]),
    }
    # forward values by the defining sum (psi = 2 mod 17, psi = 4 mod 257)
    fwd = np.zeros((1, 4, 1, 2), dtype=np.int64)
    for t, (q, psi) in enumerate([(17, 2), (257, 4)]):
        col = [a[0][i][0][t] for i in range(4)]
        for u in range(4):
            fwd[0, u, 0, t] = sum(col[i] * pow(psi, (2 * u + 1) * i, q) for i in range(4)) % q
    add("ntt_known_answer", "tests/test_ntt.cpp:7-53 (forward values by the defining sum)", "ntt", ntt_args, val(fwd))
    add("intt_roundtrip_input", "tests/test_ntt.cpp:43-50", "intt",
        {"a": T(fwd), "p": T([17, 257]), "perm": T([0, 2, 1, 3]),
         "twiddles": T([[1, 13, 9, 15], [1, 241, 193, 253]]), "m_inv": T([13, 193])}, val(a))

    # ---------------- g_decomposition (tests/test_g_decomposition.cpp)
    add("gdecomp_scalar_values", "tests/test_g_decomposition.cpp:7-25", "g_decomposition",
        {"a": T([0, 1, 2, 3]), "result_shape": [4, 2], "power": 2, "base_bits": 1},
        {**val([[0, 0], [1, 0], [0, 1], [1, 1]]), "overflow": 0})
    add("gdecomp_zero_input", "tests/test_g_decomposition.cpp:27-35", "g_decomposition",
        {"a": T([0] * 5), "result_shape": [5, 4], "power": 4, "base_bits": 2}, {**val(np.zeros((5, 4))), "overflow": 0})
    add("gdecomp_multidim", "tests/test_g_decomposition.cpp:37-60", "g_decomposition",
        {"a": T([[[5, 12], [3, 1]], [[8, 7], [9, 2]]]), "result_shape": [2, 2, 2, 3], "power": 3, "base_bits": 2},
        {**val([[[[1, 1, 0], [0, 3, 0]], [[3, 0, 0], [1, 0, 0]]], [[[0, 2, 0], [3, 1, 0]], [[1, 2, 0], [2, 0, 0]]]]),
         "overflow": 0})
    add("gdecomp_no_overflow_255", "tests/test_g_decomposition.cpp:62-72", "g_decomposition",
        {"a": T([255]), "result_shape": [1, 3], "power": 3, "base_bits": 3}, {**val([[7, 7, 3]]), "overflow": 0})
    add("gdecomp_shape_mismatch", "tests/test_g_decomposition.cpp:74-83", "g_decomposition",
        {"a": T([10, 20]), "result_shape": [3, 2], "power": 2, "base_bits": 2}, err_inv)

    # ---------------- decompose -> modmul by 2^i -> axis_modsum (tests/test_decomp_reconstruct.cpp:7-38)
    add("decomp_reconstruct", "tests/test_decomp_reconstruct.cpp:7-38", "decomp_reconstruct",
        {"a": T([51, 29, 63]), "power": 6, "base_bits": 1, "basis": T([1, 2, 4, 8, 16, 32]),
         "p6": T([1024] * 6), "p1": T([1024])}, val([51, 29, 63]))

    # ---------------- axis_modsum (tests/test_axis_modsum.cpp)
    a3 = np.array([[[1, 2, 3, 4], [5, 6, 7, 8], [9, 10, 11, 12]], [[13, 14, 15, 16], [17, 18, 19, 20], [21, 22, 23, 24]]])
    p4 = np.array([11, 13, 17, 19])
    add("axis_modsum_3d_axis1", "tests/test_axis_modsum.cpp:7-27", "axis_modsum",
        {"a": T(a3), "p": T(p4), "result_shape": [2, 4], "axis": 1}, val(a3.sum(1) % p4))
    a2 = np.array([[1, 2, 3, 4], [4, 5, 6, 7], [7, 8, 9, 10]])
    p2 = np.array([5, 7, 11, 13])
    add("axis_modsum_first_axis", "tests/test_axis_modsum.cpp:29-48", "axis_modsum",
        {"a": T(a2), "p": T(p2), "result_shape": [4], "axis": 0}, val(a2.sum(0) % p2))
    a4 = np.arange(24).reshape(2, 2, 2, 3)
    p3 = np.array([7, 11, 13])
    add("axis_modsum_high_dim", "tests/test_axis_modsum.cpp:50-64", "axis_modsum",
        {"a": T(a4), "p": T(p3), "result_shape": [2, 2, 3], "axis": 2}, val(a4.sum(2) % p3))
    ar = rnd(rng, 0, 10, (2, 2, 2))
    for ax in (-1, 3):
        add(f"axis_modsum_invalid_axis_{ax}", "tests/test_axis_modsum.cpp:66-76", "axis_modsum",
            {"a": T(ar), "p": T([7, 11]), "result_shape": [2, 2], "axis": ax}, err_inv)
    add("axis_modsum_p_mismatch", "tests/test_axis_modsum.cpp:78-85", "axis_modsum",
        {"a": T(ar), "p": T([7, 11, 13]), "result_shape": [2, 2], "axis": 2}, err_inv)

    # ---------------- modmul / modsum (tests/test_modop.cpp:101-210)
    x22 = [[1, 2], [3, 4]]
    add("modsum_tcc", "tests/test_modop.cpp:102-107", "modsum",
        {"a": T(x22), "b": 5, "p": 6, "result_shape": [2, 2]}, val((np.array(x22) + 5) % 6))
    add("modmul_tcc", "tests/test_modop.cpp:109-114", "modmul",
        {"a": T(x22), "b": 5, "p": 6, "result_shape": [2, 2]}, val((np.array(x22) * 5) % 6))
    add("modsum_tcc_scalars", "tests/test_modop.cpp:116-121", "modsum",
        {"a": T([[10, 20]]), "b": 3, "p": 7, "result_shape": [1, 2]}, val((np.array([[10, 20]]) + 3) % 7))
    add("modsum_tct", "tests/test_modop.cpp:124-129", "modsum",
        {"a": T(x22), "b": 5, "p": T([6, 7]), "result_shape": [2, 2]}, val((np.array(x22) + 5) % np.array([6, 7])))
    add("modmul_tct", "tests/test_modop.cpp:131-136", "modmul",
        {"a": T(x22), "b": 5, "p": T([6, 7]), "result_shape": [2, 2]}, val((np.array(x22) * 5) % np.array([6, 7])))
    add("modsum_tct_scalar_b", "tests/test_modop.cpp:138-143", "modsum",
        {"a": T([[10, 20]]), "b": 3, "p": T([4, 5]), "result_shape": [1, 2]},
        val((np.array([[10, 20]]) + 3) % np.array([4, 5])))
    b22 = [[5, 6], [7, 8]]
    add("modsum_ttc", "tests/test_modop.cpp:146-151", "modsum",
        {"a": T(x22), "b": T(b22), "p": 9, "result_shape": [2, 2]}, val((np.array(x22) + np.array(b22)) % 9))
    add("modmul_ttc", "tests/test_modop.cpp:153-158", "modmul",
        {"a": T(x22), "b": T(b22), "p": 9, "result_shape": [2, 2]}, val((np.array(x22) * np.array(b22)) % 9))
    add("modsum_ttc_scalar_p", "tests/test_modop.cpp:160-165", "modsum",
        {"a": T([[10, 20]]), "b": T([[1, 2]]), "p": 7, "result_shape": [1, 2]},
        val((np.array([[10, 20]]) + np.array([[1, 2]])) % 7))
    ra, rb, rp = rnd(rng, 1, 100, (2, 3, 4)), rnd(rng, 1, 100, (1, 3, 4)), rnd(rng, 100, 200, (4,))
    add("modsum_ttt_broadcast_3d", "tests/test_modop.cpp:169-174", "modsum",
        {"a": T(ra), "b": T(rb), "p": T(rp), "result_shape": [2, 3, 4]}, val((ra + rb) % rp))
    add("modmul_ttt_broadcast_rightmost", "tests/test_modop.cpp:176-181", "modmul",
        {"a": T([[1], [2]]), "b": T([10, 20, 30]), "p": T([11, 17, 23]), "result_shape": [2, 3]},
        val((np.array([[1], [2]]) * np.array([10, 20, 30])) % np.array([11, 17, 23])))
    add("modsum_ttt_broadcast_rightmost", "tests/test_modop.cpp:183-188", "modsum",
        {"a": T([[5], [8]]), "b": T([1, 2, 3]), "p": T([6, 7, 8]), "result_shape": [2, 3]},
        val((np.array([[5], [8]]) + np.array([1, 2, 3])) % np.array([6, 7, 8])))
    add("modsum_incompatible_shapes", "tests/test_modop.cpp:190-199", "modsum",
        {"a": T(rnd(rng, 0, 10, (2, 3))), "b": T(rnd(rng, 0, 10, (2, 4))), "p": T(rnd(rng, 100, 200, (3,))),
         "result_shape": [2, 3]}, err_inv)
    add("modsum_incorrect_p_shape", "tests/test_modop.cpp:201-210", "modsum",
        {"a": T(rnd(rng, 0, 10, (2, 3))), "b": T(rnd(rng, 0, 10, (2, 3))), "p": T(rnd(rng, 100, 200, (4, 1))),
         "result_shape": [2, 3]}, err_inv)
    # tests/test_noncontiguous.cpp:6-32
    base = np.arange(18).reshape(2, 3, 3, 1)
    add("modmul_ttt_noncontiguous", "tests/test_noncontiguous.cpp:6-32", "modmul",
        {"a": T(base), "b": T(base, view=[["transpose", 1, 2]]), "p": T([100]), "result_shape": [2, 3, 3, 1]},
        val((base * np.swapaxes(base, 1, 2)) % 100))

    # ---------------- mod_tt / mod_tc / mod_ct (tests/test_modop.cpp:224-486)
    A = np.array([[5, 7, 9], [10, 12, 14]])
    B = np.array([[3, 5, 7], [4, 6, 8]])
    add("mod_tt_basic", "tests/test_modop.cpp:224-240", "mod",
        {"a": T(A, "int64"), "b": T(B, "int64"), "result_shape": [2, 3]}, val(A % B, "int64"))
    add("mod_tt_singleton", "tests/test_modop.cpp:243-252", "mod",
        {"a": T([[[7]]], "int64"), "b": T([[[3]]], "int64"), "result_shape": [1, 1, 1]}, val([[[1]]], "int64"))
    ia, ib = rnd(rng, 1, 100, (2, 2, 2, 2)), rnd(rng, 1, 100, (2, 2, 2, 2))
    add("mod_tt_int32", "tests/test_modop.cpp:258-267", "mod",
        {"a": T(ia), "b": T(ib), "result_shape": [2, 2, 2, 2]}, val(ia % ib))
    bA, bB = np.arange(1, 25).reshape(2, 3, 4), np.arange(25, 49).reshape(2, 3, 4)
    add("mod_tt_noncontiguous", "tests/test_modop.cpp:269-287", "mod",
        {"a": T(bA, "int64", view=[["transpose", 0, 2]]), "b": T(bB, "int64", view=[["permute", [2, 1, 0]]]),
         "result_shape": [4, 3, 2]}, val(np.swapaxes(bA, 0, 2) % np.transpose(bB, (2, 1, 0)), "int64"))
    add("mod_tt_null_a", "tests/test_modop.cpp:293-299", "mod",
        {"a": None, "b": T([1, 2, 3], "int64"), "result_shape": [3]}, err_inv)
    add("mod_tt_null_b", "tests/test_modop.cpp:293-299", "mod",
        {"a": T([1, 2, 3], "int64"), "b": None, "result_shape": [3]}, err_inv)
    add("mod_tt_shape_mismatch", "tests/test_modop.cpp:301-308", "mod",
        {"a": T(rnd(rng, 0, 10, (2, 3)), "int64"), "b": T(rnd(rng, 0, 10, (2, 2)), "int64"), "result_shape": [2, 3]},
        err_inv)
    add("mod_tc_basic", "tests/test_modop.cpp:323-338", "mod",
        {"a": T(A, "int64"), "b": 6, "result_shape": [2, 3]}, val(A % 6, "int64"))
    add("mod_tc_singleton", "tests/test_modop.cpp:340-346", "mod",
        {"a": T([[7]], "int64"), "b": 3, "result_shape": [1, 1]}, val([[1]], "int64"))
    ic = rnd(rng, 1, 50, (2, 2, 3, 2))
    add("mod_tc_int32", "tests/test_modop.cpp:352-358", "mod",
        {"a": T(ic), "b": 7, "result_shape": [2, 2, 3, 2]}, val(ic % 7))
    b3 = np.arange(1, 13).reshape(2, 2, 3)
    add("mod_tc_noncontiguous", "tests/test_modop.cpp:360-367", "mod",
        {"a": T(b3, "int64", view=[["transpose", 1, 2]]), "b": 4, "result_shape": [2, 3, 2]},
        val(np.swapaxes(b3, 1, 2) % 4, "int64"))
    add("mod_tc_null", "tests/test_modop.cpp:373-378", "mod",
        {"a": None, "b": 5, "result_shape": [3]}, err_inv)
    add("mod_tc_shape_mismatch", "tests/test_modop.cpp:380-386", "mod",
        {"a": T(rnd(rng, 0, 10, (2, 3)), "int64"), "b": 5, "result_shape": [3, 2]}, err_inv)
    add("mod_ct_basic", "tests/test_modop.cpp:401-417", "mod",
        {"a": 6, "b": T(B, "int64"), "result_shape": [2, 3]}, val(6 % B, "int64"))
    z = np.array([[1, 2], [3, 4]])
    add("mod_ct_zero_numerator", "tests/test_modop.cpp:419-435", "mod",
        {"a": 0, "b": T(z, "int64"), "result_shape": [2, 2]}, val(0 % z, "int64"))
    add("mod_ct_singleton", "tests/test_modop.cpp:437-444", "mod",
        {"a": 8, "b": T([[3]], "int64"), "result_shape": [1, 1]}, val([[2]], "int64"))
    idd = rnd(rng, 1, 20, (2, 3, 2, 2))
    add("mod_ct_int32", "tests/test_modop.cpp:450-457", "mod",
        {"a": 9, "b": T(idd), "result_shape": [2, 3, 2, 2]}, val(9 % idd))
    b4 = np.arange(1, 19).reshape(2, 3, 3)
    add("mod_ct_noncontiguous", "tests/test_modop.cpp:459-467", "mod",
        {"a": 4, "b": T(b4, "int64", view=[["transpose", 0, 2]]), "result_shape": [3, 3, 2]},
        val(4 % np.swapaxes(b4, 0, 2), "int64"))
    add("mod_ct_null", "tests/test_modop.cpp:474-479", "mod",
        {"a": 5, "b": None, "result_shape": [3]}, err_inv)
    add("mod_ct_shape_mismatch", "tests/test_modop.cpp:481-486", "mod",
        {"a": 5, "b": T(rnd(rng, 0, 10, (2, 3)), "int64"), "result_shape": [3, 2]}, err_inv)

    # ---------------- permute (tests/test_permute.cpp)
    def perm_expected(a, perms, ew, pa):
        moved = np.moveaxis(a, (ew, pa), (-2, -1))
        pe = np.broadcast_to(perms.reshape((1,) * (moved.ndim - 2) + perms.shape), moved.shape)
        out = np.take_along_axis(moved, pe, axis=-1)
        return np.moveaxis(out, (-2, -1), (ew, pa))

    for name, cite, shape, perms, ew, pa in [
        ("permute_rank4_0_1", "tests/test_permute.cpp:46-50", (2, 3, 4, 5), [[2, 1, 0], [0, 2, 1]], 0, 1),
        ("permute_rank5_0_3", "tests/test_permute.cpp:52-56", (2, 2, 2, 3, 5), [[1, 0, 2], [2, 1, 0]], 0, 3),
        ("permute_inner_dims", "tests/test_permute.cpp:58-62", (3, 4, 5), [[3, 2, 1, 0], [0, 2, 3, 1], [1, 0, 3, 2]], 0, 1),
        ("permute_dim1_ew0", "tests/test_permute.cpp:86-97", (5, 3, 7),
         [[2, 1, 0], [0, 2, 1], [1, 0, 2], [2, 0, 1], [0, 1, 2]], 0, 1),
        ("permute_dim1_ew2", "tests/test_permute.cpp:99-107", (3, 4, 2), [[1, 2, 0, 3], [3, 0, 1, 2]], 2, 1),
    ]:
        a = rnd(rng, 0, 100, shape)
        pm = np.array(perms)
        add(name, cite, "permute", {"a": T(a), "perms": T(pm), "elementwise_axis": ew, "perm_axis": pa},
            val(perm_expected(a, pm, ew, pa)))
    ai = rnd(rng, 0, 10, (2, 3, 4))
    pi = rnd(rng, 0, 3, (2, 3))
    for ew, pa in [(3, 1), (1, 1), (0, 3)]:
        add(f"permute_invalid_axes_{ew}_{pa}", "tests/test_permute.cpp:64-74", "permute",
            {"a": T(ai), "perms": T(pi), "elementwise_axis": ew, "perm_axis": pa}, err_inv)
    add("permute_out_of_bounds", "tests/test_permute.cpp:76-84", "permute",
        {"a": T(ai), "perms": T([[0, 1, 3], [1, 2, 0]]), "elementwise_axis": 0, "perm_axis": 1}, err_oor)

    # ---------------- memory / views (tests/test_memory_ops.cpp, test_reshape.cpp, test_contiguous.cpp)
    add("squeeze_singleton", "tests/test_memory_ops.cpp:7-16", "squeeze",
        {"a": T([[[1, 2, 3]]], "int64"), "axis": 1}, val([[1, 2, 3]], "int64"))
    add("squeeze_not_one", "tests/test_memory_ops.cpp:18-23", "squeeze",
        {"a": T([[1, 2, 3]], "int64"), "axis": 1}, err_inv)
    add("unsqueeze_axis1", "tests/test_memory_ops.cpp:25-34", "unsqueeze",
        {"a": T([[1, 2], [3, 4]]), "axis": 1}, val([[[1, 2]], [[3, 4]]]))
    add("unsqueeze_negative", "tests/test_memory_ops.cpp:36-45", "unsqueeze",
        {"a": T([[1, 2], [3, 4]]), "axis": -1}, val([[[1], [2]], [[3], [4]]]))
    ru = rnd(rng, 0, 10, (2, 3))
    for ax in (4, -4):
        add(f"unsqueeze_out_of_range_{ax}", "tests/test_memory_ops.cpp:47-53", "unsqueeze",
            {"a": T(ru, "int64"), "axis": ax}, err_inv)
    c = [[[1, 2, 3], [4, 5, 6]], [[7, 8, 9], [10, 11, 12]]]
    add("reshape_6x2", "tests/test_reshape.cpp:7-20", "reshape", {"a": T(c), "shape": [6, 2]},
        val(np.array(c).reshape(6, 2)))
    add("reshape_3x4", "tests/test_reshape.cpp:21-33", "reshape", {"a": T(c), "shape": [3, 4]},
        val(np.array(c).reshape(3, 4)))
    t = np.arange(12).reshape(3, 4)
    add("make_contiguous_transposed", "tests/test_contiguous.cpp:6-12", "make_contiguous",
        {"a": T(t, view=[["transpose", 0, 1]])}, val(t.T))
    add("make_contiguous_same_pointer", "tests/test_contiguous.cpp:14-20", "make_contiguous",
        {"a": T(rnd(rng, 0, 60000, (5, 6)), "int64")}, {"same_object": True})
    return cases


def ntt_closed_form_cases():
    import sys

    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from heal_amd import params as P

    rng = random.Random(7)
    out = []
    for m in (16, 64):
        for bits in (30, 60):
            primes = P.ntt_primes(m, [bits, bits])
            prm = P.NttParams(m, primes)
            a = [[rng.randrange(0, q) for q in primes] for _ in range(m)]  # [m][k]
            fwd = [[0] * len(primes) for _ in range(m)]
            for t, (q, psi) in enumerate(zip(primes, prm.psi)):
                for u in range(m):
                    fwd[u][t] = sum(a[i][t] * pow(psi, (2 * u + 1) * i, q) for i in range(m)) % q
            out.append({
                "m": m, "bits": bits, "primes": [str(q) for q in primes], "psi": [str(g) for g in prm.psi],
                "a": [[str(x) for x in row] for row in a], "expected": [[str(x) for x in row] for row in fwd],
            })
    return out


def ntt_signed_cases():
    """Length-2 NTT/INTT vectors worked by hand from the reference's loops, with
    C++'s truncating % (example_impl/ntt_impl.cpp:101-108 forward, :151-157
    inverse, :165 scaling).  q = 17, twiddles [1, 4] (psi = 4 is a primitive
    4th root), inverse twiddles [1, 13] (13 = 4^-1), m_inv = 9 (= 2^-1).
    Forward:  x0 = (a0 + v) % q, x1 = (a0 + q - v) % q with v = (a1 * 4) % q,
              then out[u] = x[perm[u]].
    Inverse:  x[perm[u]] = a[u]; x0 = (x0 + x1) % q, x1 = ((x0 + q - x1) * 13) % q,
              then each times 9, % q.
    Worked examples (the sign of every remainder is the sign of its dividend):
      fwd [-1, 0]:   v = 0;  (-1) % 17 = -1;  16 % 17 = 16                  -> [-1, 16]
      fwd [3, -2]:   v = -8; (-5) % 17 = -5;  28 % 17 = 11                  -> [-5, 11]
      fwd [3, -2], perm [1, 0]                                              -> [11, -5]
      fwd [-20, -3]: v = -12; (-32) % 17 = -15; 9 % 17 = 9                  -> [-15, 9]
      inv [-5, 3]:   -2; 9 * 13 = 117 -> 15; -18 % 17 = -1; 135 % 17 = 16  -> [-1, 16]
      inv [2, 30]:   32 -> 15; -11 * 13 = -143 -> -7; 135 -> 16; -63 -> -12 -> [16, -12]
      inv [2, 30], perm [1, 0]: x = [30, 2]: 15; 45 * 13 = 585 -> 7; 16; 63 -> 12 -> [16, 12]
      inv [-20, -3]: -23 -> -6; 0; -54 -> -3; 0                             -> [-3, 0]
    Every case is stored for int32 and int64 as an [l=1, m=2, r=1, k=1] tensor."""
    base = {"p": [17], "tw": [[1, 4]], "itw": [[1, 13]], "m_inv": [9]}
    rows = [
        ("fwd", [-1, 0], [0, 1], [-1, 16]),
        ("fwd", [3, -2], [0, 1], [-5, 11]),
        ("fwd", [3, -2], [1, 0], [11, -5]),
        ("fwd", [-20, -3], [0, 1], [-15, 9]),
        ("inv", [-5, 3], [0, 1], [-1, 16]),
        ("inv", [2, 30], [0, 1], [16, -12]),
        ("inv", [2, 30], [1, 0], [16, 12]),
        ("inv", [-20, -3], [0, 1], [-3, 0]),
    ]
    cases = []
    for dtype in ("int32", "int64"):
        for i, (kind, a, perm, want) in enumerate(rows):
            c = dict(base)
            c.update({"name": f"signed_{kind}_{i}_{dtype}", "kind": kind, "dtype": dtype, "a": a, "perm": perm,
                      "want": want})
            cases.append(c)
    return cases


if __name__ == "__main__":
    with open(os.path.join(HERE, "reference_kat.json"), "w") as f:
        json.dump(reference_cases(), f, indent=0)
    with open(os.path.join(HERE, "ntt_closed_form.json"), "w") as f:
        json.dump(ntt_closed_form_cases(), f, indent=0)
    with open(os.path.join(HERE, "ntt_signed_kat.json"), "w") as f:
        json.dump(ntt_signed_cases(), f, indent=0)
    print("wrote", os.listdir(HERE))
