"""CPU: the host Fp / G1 Jacobian arithmetic of the MSM tail
(zkwasm-prover_amd/csrc/host_ec.h), compiled with g++.

Field operations against python integers in Montgomery form (R = 2^256) over
edge values (0, 1, p - 1, p - 2, values whose Montgomery form is p - 1, carries
across every word) and seeded random ones; curve operations against the
oracle's G1 arithmetic on points with non-trivial Z, including the doubling,
cancellation and identity cases and the XYZZ conversion the tail applies to
the device's masked sums, and the MSM tail (one Horner pass over the
masked sums' exponents) against the oracle's multiple of G for the weighted
sum of the masked sums' discrete logarithms.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
P = 21888242871839275222246405745257275088696311157297823662689037894645226208583
R256 = 1 << 256
RINV = pow(R256, -1, P)


@pytest.fixture(scope="module")
def he(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("he") / "libhe.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I",
                    os.path.join(ROOT, "zkwasm-prover_amd", "csrc"),
                    os.path.join(HERE, "host_ec", "check.cpp"), "-o", out], check=True)
    L = ctypes.CDLL(out)
    for f in ("he_add", "he_sub", "he_mul", "he_jac_add"):
        getattr(L, f).argtypes = [ctypes.c_void_p] * 3
    for f in ("he_inv", "he_jac_dbl", "he_jac_normalize", "he_from_xyzz"):
        getattr(L, f).argtypes = [ctypes.c_void_p] * 2
    L.he_msm_tail.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_uint32] * 4
    L.he_msm_tail_split.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_uint32] * 4 + [ctypes.c_int]
    return L


def limbs(v):
    return np.array([(v >> (64 * i)) & (2**64 - 1) for i in range(4)], dtype=np.uint64)


def to_int(a):
    return sum(int(a[i]) << (64 * i) for i in range(4))


def mont(v):
    return limbs(v * R256 % P)


def unmont(a):
    return to_int(a) * RINV % P


def _values():
    edge = [0, 1, 2, P - 1, P - 2, (P - 1) // 2, 2**64 - 1, 2**128 + 5, 2**192 - 1, 2**253,
            (P - 1) * RINV % P, (P - 2) * RINV % P, 1 * RINV % P]
    rng = np.random.default_rng(31)
    rand = [int.from_bytes(rng.bytes(32), "little") % P for _ in range(40)]
    return edge + rand


def _call2(fn, a, b):
    r = np.zeros(4, dtype=np.uint64)
    fn(r.ctypes.data, a.ctypes.data, b.ctypes.data)
    return r


@pytest.mark.parametrize("op", ["add", "sub", "mul"])
def test_field_ops(he, op):
    vals = _values()
    fn = getattr(he, "he_" + op)
    exp = {"add": lambda x, y: (x + y) % P, "sub": lambda x, y: (x - y) % P, "mul": lambda x, y: x * y % P}[op]
    for i, x in enumerate(vals):
        for y in (vals[0], vals[3], vals[i], vals[-1 - i], vals[(7 * i + 3) % len(vals)]):
            r = _call2(fn, mont(x), mont(y))
            assert to_int(r) < P, "not reduced"
            assert unmont(r) == exp(x, y), (op, x, y)


def test_field_inverse(he):
    for x in _values():
        r = np.zeros(4, dtype=np.uint64)
        a = mont(x)
        he.he_inv(r.ctypes.data, a.ctypes.data)
        assert unmont(r) == (pow(x, -1, P) if x else 0)


def _jac(zo, k, lam):
    """k G as Jacobian (lam^2 x, lam^3 y, lam z) (Montgomery limbs, 12 u64)"""
    aff = zo.g1_to_affine(zo.g1_mul(zo.g1_generator(), zo.to_mont([k])))
    x, y = zo.g1a_to_ints(aff)
    return np.concatenate([mont(x * lam * lam % P), mont(y * pow(lam, 3, P) % P), mont(lam)]), (x, y)


def _affine(he, j):
    r = np.zeros(12, dtype=np.uint64)
    he.he_jac_normalize(r.ctypes.data, np.ascontiguousarray(j).ctypes.data)
    if to_int(r[8:]) == 0:
        return None
    assert unmont(r[8:]) == 1
    return unmont(r[:4]), unmont(r[4:8])


def _oracle_affine(zo, k):
    if k % 21888242871839275222246405745257275088548364400416034343698204186575808495617 == 0:
        return None
    return zo.g1a_to_ints(zo.g1_to_affine(zo.g1_mul(zo.g1_generator(), zo.to_mont([k]))))


@pytest.mark.parametrize("ka,kb", [(3, 5), (12345, 2**200 + 7), (7, 7), (9, -9), (11, 0)])
def test_jacobian_add(he, zo, ka, kb):
    r_ = 21888242871839275222246405745257275088548364400416034343698204186575808495617
    a, _ = _jac(zo, ka % r_, 0x1234567)
    if kb % r_ == 0:
        b = np.concatenate([mont(0), mont(1), mont(0)])  # the identity (z = 0)
    else:
        b, _ = _jac(zo, kb % r_, 0xfedcba98765)
    out = np.zeros(12, dtype=np.uint64)
    he.he_jac_add(out.ctypes.data, a.ctypes.data, b.ctypes.data)
    assert _affine(he, out) == _oracle_affine(zo, (ka + kb) % r_)
    he.he_jac_add(out.ctypes.data, b.ctypes.data, a.ctypes.data)
    assert _affine(he, out) == _oracle_affine(zo, (ka + kb) % r_)


@pytest.mark.parametrize("k", [1, 2, 99991, 2**250 + 3])
def test_jacobian_double(he, zo, k):
    a, _ = _jac(zo, k, 0xabcdef0123)
    out = np.zeros(12, dtype=np.uint64)
    he.he_jac_dbl(out.ctypes.data, a.ctypes.data)
    assert _affine(he, out) == _oracle_affine(zo, 2 * k)
    ident = np.concatenate([mont(0), mont(1), mont(0)])
    he.he_jac_dbl(out.ctypes.data, ident.ctypes.data)
    assert _affine(he, out) is None


def test_from_xyzz(he, zo):
    _, (x, y) = _jac(zo, 424242, 1)
    lam = 0x5555aaaa1234
    zz, zzz = lam * lam % P, pow(lam, 3, P)
    xyzz = np.concatenate([mont(x * zz % P), mont(y * zzz % P), mont(zz), mont(zzz)])
    out = np.zeros(12, dtype=np.uint64)
    he.he_from_xyzz(out.ctypes.data, xyzz.ctypes.data)
    assert _affine(he, out) == (x, y)
    he.he_from_xyzz(out.ctypes.data, np.concatenate([mont(1), mont(1), mont(0), mont(0)]).ctypes.data)
    assert _affine(he, out) is None


R_ORDER = 21888242871839275222246405745257275088548364400416034343698204186575808495617


def _xyzz(zo, k, lam):
    """k G in XYZZ form (lam^2 x, lam^3 y, lam^2, lam^3); k = 0 -> ZZ = 0"""
    if k % R_ORDER == 0:
        return np.concatenate([mont(1), mont(1), mont(0), mont(0)])
    _, (x, y) = _jac(zo, k % R_ORDER, 1)
    zz, zzz = lam * lam % P, pow(lam, 3, P)
    return np.concatenate([mont(x * zz % P), mont(y * zzz % P), mont(zz), mont(zzz)])


def _tail(he, zo, ks, c, nb, l1=3, seed=0, split=None):
    """ks[w][b], b <= nb: discrete logs of the masked sums; returns (got,
    expected); split: the exponent at which the sum is cut in two parts"""
    rng = np.random.default_rng(seed)
    W = len(ks)
    q = np.concatenate([_xyzz(zo, ks[w][b], int(rng.integers(2, 2**62))) for w in range(W) for b in range(nb + 1)])
    out = np.zeros(12, dtype=np.uint64)
    if split is None:
        he.he_msm_tail(out.ctypes.data, np.ascontiguousarray(q).ctypes.data, c, W, nb, l1)
    else:
        he.he_msm_tail_split(out.ctypes.data, np.ascontiguousarray(q).ctypes.data, c, W, nb, l1, split)
    k = sum((ks[w][nb] + sum(ks[w][b] << (l1 + b) for b in range(nb))) << (c * w) for w in range(W)) % R_ORDER
    return _affine(he, out), _oracle_affine(zo, k)


@pytest.mark.parametrize("c,W,nb", [(17, 15, 13), (16, 16, 12), (5, 51, 1), (4, 64, 0), (2, 128, 0), (13, 20, 9)])
def test_msm_tail(he, zo, c, W, nb):
    """seeded random masked sums at the slice shapes of windows 2..17 (nb =
    c - 1 - 3 bits, 0 when a window has at most 8 buckets), with identities"""
    rng = np.random.default_rng(c * 1000 + W)
    ks = [[int.from_bytes(rng.bytes(32), "little") % R_ORDER for _ in range(nb + 1)] for _ in range(W)]
    ks[W // 2][0] = 0
    ks[0][nb] = 0
    ks[W - 1] = [0] * (nb + 1)
    got, exp = _tail(he, zo, ks, c, nb, seed=c)
    assert got == exp
    top = c * (W - 1) + (3 + nb if nb else 1) - 1
    for lo in (1, top * 11 // 20, top):
        assert _tail(he, zo, ks, c, nb, seed=c, split=lo) == (exp, exp)


def test_msm_tail_doubling_cancel_identity(he, zo):
    """a term equal to the running sum (doubling inside the addition), its
    negation (cancellation to the identity), and all-identity sums"""
    # W = 2, nb = 0, c = 2: acc = 4 Q[1][0] when Q[0][0] is added
    got, exp = _tail(he, zo, [[5], [5 * pow(4, -1, R_ORDER)]], 2, 0)
    assert got == exp == _oracle_affine(zo, 10)
    got, exp = _tail(he, zo, [[R_ORDER - 8], [2]], 2, 0)
    assert got is None and exp is None
    got, exp = _tail(he, zo, [[0, 0, 0], [0, 0, 0]], 9, 2)
    assert got is None and exp is None
    # two windows whose term ranges overlap (l1 + nb > c): several terms per exponent
    got, exp = _tail(he, zo, [[3, 1, 4, 1, 5], [9, 2, 6, 5, 3], [5, 8, 9, 7, 9]], 5, 4)
    assert got == exp
