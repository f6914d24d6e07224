"""CPU: the host Fr arithmetic the prover core uses for the opening scalars
(zkwasm-prover_amd/csrc/host_fr.h) against the oracle and python integers,
including the edge values 0, 1, r - 1 and carries across every limb."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
R = 21888242871839275222246405745257275088548364400416034343698204186575808495617


@pytest.fixture(scope="module")
def hf(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("hf") / "libhf.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", os.path.join(ROOT, "zkwasm-prover_amd", "csrc"),
                    os.path.join(HERE, "host_fr", "check.cpp"), "-o", out], check=True)
    return ctypes.CDLL(out)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _vals(zo):
    edge = zo.to_mont([0, 1, 2, R - 1, R - 2, 2**64 - 1, 2**128 + 5, 2**192 - 1, (R - 1) // 2])
    return np.concatenate([edge, zo.rand_fr(40, 99)])


@pytest.mark.parametrize("op", ["add", "sub", "mul"])
def test_host_fr_binary(zo, hf, op):
    v = _vals(zo)
    ref = {"add": zo.fr_add, "sub": zo.fr_sub, "mul": zo.fr_mul}[op]
    fn = getattr(hf, "hf_" + op)
    for i in range(len(v)):
        for jdx in (0, 3, i, len(v) - 1 - i):
            a, b = v[i:i + 1].copy(), v[jdx:jdx + 1].copy()
            r = zo.fr_zeros(1)
            fn(_p(r), _p(a), _p(b))
            assert (r == ref(a, b)).all(), (op, i, jdx)


def test_host_fr_inverse_and_small(zo, hf):
    v = _vals(zo)
    for i in range(len(v)):
        a = v[i:i + 1].copy()
        r = zo.fr_zeros(1)
        hf.hf_inv(_p(r), _p(a))
        assert (r == zo.fr_inv(a)).all()
        if a.any():
            assert zo.from_mont(zo.fr_mul(r, a))[0] == 1
    for x in (0, 1, 7, 2**63, 2**64 - 1):
        r = zo.fr_zeros(1)
        hf.hf_from_u64(_p(r), ctypes.c_ulonglong(x))
        assert zo.from_mont(r)[0] == x % R
