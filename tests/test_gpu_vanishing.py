"""GPU parity: the vanishing construction (src/eval_h.rs:464-583) through the
host mirror zkp_vanishing_commit / zkp_vanishing_pieces over the C-ABI
(field_mul_zip, ntt, extended_prepare, MSMs, field_op), against the oracle's
composition of the same steps (itself pinned by the coset round trip and
commitment linearity on the CPU).  Commitments, coefficients and h_pieces
must match bit for bit, with the reference driver's MSMs and the batched MSM.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

R = 21888242871839275222246405745257275088548364400416034343698204186575808495617


@pytest.mark.parametrize("log_n,log_ext,q,batched", [(10, 12, 3, 0), (10, 12, 3, 1), (12, 14, 4, 1)])
def test_vanishing_matches_oracle(gpu, zo, log_n, log_ext, q, batched):
    from zkprover import device as dv
    from zkprover._lib import check, lib
    n, N = 1 << log_n, 1 << log_ext
    seed = 60 + log_n + batched
    h = zo.rand_fr(N, seed)
    t = zo.rand_fr_nonzero(N // n, seed + 1)
    w_inv = zo.fr_inv(zo.root_of_unity(log_ext))
    divisor = zo.fr_inv(zo.fr_from_u64(N))
    zeta = zo.to_mont([pow(7, (R - 1) // 3, R)])
    powers = np.concatenate([zo.fr_mul(zeta, zeta), zeta])  # [g_coset_inv, g_coset]
    g = zo.gen_bases_random(n, seed + 2)
    xn = zo.rand_fr(1, seed + 3)
    exp_coeffs, exp_commits, exp_pieces = zo.vanishing_construct(h, t, log_n, log_ext, w_inv, divisor, powers,
                                                                 g, q, xn)

    init = np.zeros((N, 4), dtype=np.uint64)
    init[0] = zo.fr_from_u64(1)[0]
    init[1] = w_inv[0]
    om = dv.to_dev(init)
    check(lib().expand_omega_buffer(ctypes.c_void_p(om.data_ptr()), log_ext, 1), "expand_omega_buffer")
    dh, dtmp = dv.to_dev(h), dv.empty(N)
    dt, ddiv, dpow, dg, dxn = dv.to_dev(t), dv.to_dev(divisor), dv.to_dev(powers), dv.to_dev(g.reshape(-1, 4)), \
        dv.to_dev(xn)
    commits = np.zeros((q, 8), dtype=np.uint64)
    coeffs = ctypes.c_void_p()
    f = lib().zkp_vanishing_commit
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int] + [ctypes.c_void_p] * 3 + [ctypes.c_int] * 3 + \
        [ctypes.c_void_p] * 3 + [ctypes.c_int]
    check(f(dh.data_ptr(), dtmp.data_ptr(), dt.data_ptr(), len(t), om.data_ptr(), ddiv.data_ptr(), dpow.data_ptr(),
            log_n, log_ext, q, dg.data_ptr(), commits.ctypes.data, ctypes.byref(coeffs), batched),
          "zkp_vanishing_commit")
    assert coeffs.value in (dh.data_ptr(), dtmp.data_ptr())
    got_coeffs = dv.from_dev(dh if coeffs.value == dh.data_ptr() else dtmp)
    assert (got_coeffs == exp_coeffs).all()
    assert (commits == exp_commits).all()

    pieces = np.zeros((n, 4), dtype=np.uint64)
    f2 = lib().zkp_vanishing_pieces
    f2.restype = ctypes.c_int
    f2.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    check(f2(coeffs.value, dxn.data_ptr(), log_n, q, pieces.ctypes.data), "zkp_vanishing_pieces")
    assert (pieces == exp_pieces).all()
