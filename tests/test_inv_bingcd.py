"""CPU: the optimized binary GCD inversion (zkwasm-prover_amd/csrc/inv_bingcd.h)
compiled with g++, for the BN254 scalar and base field moduli.

For every input y, the returned v must satisfy y * v * 2^((32 - k) T) = 1
(mod M), checked with python integers: edge values (1, 2, M - 1, M - 2,
powers of two, values with long runs of equal bits, small values) and a
seeded random sample; y = 0 must report failure.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
MODS = {0: 21888242871839275222246405745257275088548364400416034343698204186575808495617,
        1: 21888242871839275222246405745257275088696311157297823662689037894645226208583}


@pytest.fixture(scope="module")
def ib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("ib") / "libib.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I",
                    os.path.join(ROOT, "zkwasm-prover_amd", "csrc"),
                    os.path.join(HERE, "inv_bingcd", "check.cpp"), "-o", out], check=True)
    L = ctypes.CDLL(out)
    L.ib_inv.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    return L


def _limbs(v):
    return np.array([(v >> (32 * i)) & 0xffffffff for i in range(8)], dtype=np.uint32)


def _int(a):
    return sum(int(a[i]) << (32 * i) for i in range(8))


def _values(M, seed, k):
    edge = [1, 2, 3, M - 1, M - 2, (M - 1) // 2, (M + 1) // 2, 2**64 - 1, 2**64, 2**128 + 1, 2**252,
            2**253 - 1, (2**253 - 1) ^ (2**100), int("5" * 70) % M, int("f" * 63, 16) % M]
    edge += [1 << s for s in range(0, 253, 7)]
    rng = np.random.default_rng(seed)
    rand = [int.from_bytes(rng.bytes(32), "little") % M for _ in range(k)]
    return [v for v in edge + rand if 0 < v < M]


@pytest.mark.parametrize("field", [0, 1])
def test_inverse_identity(ib, field):
    M = MODS[field]
    corr = pow(2, (32 - ib.ib_bits()) * ib.ib_steps(), M)
    for y in _values(M, 11 + field, 3000):
        v = np.zeros(8, dtype=np.uint32)
        ok = ib.ib_inv(v.ctypes.data, _limbs(y).ctypes.data, field)
        assert ok == 1, y
        assert _int(v) < M
        assert y * _int(v) * corr % M == 1, y


@pytest.mark.parametrize("field", [0, 1])
def test_zero_is_reported(ib, field):
    v = np.zeros(8, dtype=np.uint32)
    assert ib.ib_inv(v.ctypes.data, _limbs(0).ctypes.data, field) == 0
