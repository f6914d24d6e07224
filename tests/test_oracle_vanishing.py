"""Pins the oracle's vanishing construction (src/eval_h.rs:464-583) by
properties independent of its composition:
  * coset round trip: re-distributing the coefficients with [zeta, zeta^2]
    and running the forward extended NTT gives back h * t_evaluations, which
    fixes the iNTT, the divisor and the order of the coset powers;
  * commitment linearity: the commitment of the Horner-folded pieces equals
    sum_i xn^i * commitment_i (one scalar multiplication per piece), which fixes
    the Horner direction and the piece boundaries.
"""
import numpy as np

R = 21888242871839275222246405745257275088548364400416034343698204186575808495617


def _setup(zo, log_n, log_ext, q, seed):
    n, N = 1 << log_n, 1 << log_ext
    h = zo.rand_fr(N, seed)
    t = zo.rand_fr_nonzero(N // n, seed + 1)
    w_ext = zo.root_of_unity(log_ext)
    w_inv = zo.fr_inv(w_ext)
    divisor = zo.fr_inv(zo.fr_from_u64(N))
    zeta = zo.to_mont([pow(7, (R - 1) // 3, R)])
    zeta2 = zo.fr_mul(zeta, zeta)
    g = zo.gen_bases_random(n, seed + 2)
    xn = zo.rand_fr(1, seed + 3)
    return h, t, w_ext, w_inv, divisor, zeta, zeta2, g, xn


def test_vanishing_coset_round_trip(zo):
    log_n, log_ext, q = 8, 10, 3
    h, t, w_ext, w_inv, divisor, zeta, zeta2, g, xn = _setup(zo, log_n, log_ext, q, 40)
    coeffs, _, _ = zo.vanishing_construct(h, t, log_n, log_ext, w_inv, divisor,
                                          np.concatenate([zeta2, zeta]), g, q, xn)
    N = 1 << log_ext
    back = zo.extended_prepare(coeffs, np.concatenate([zeta, zeta2]), 1 << log_n, N, 1)
    back = zo.ntt(back, log_ext, w_ext)
    assert (back == zo.field_mul_zip(h, t)).all()
    # the iNTT alone (no coset step) would not round-trip
    plain = zo.ntt(coeffs, log_ext, w_ext)
    assert not (plain == zo.field_mul_zip(h, t)).all()


def test_vanishing_commitment_linearity(zo):
    log_n, log_ext, q = 8, 10, 3
    h, t, w_ext, w_inv, divisor, zeta, zeta2, g, xn = _setup(zo, log_n, log_ext, q, 50)
    coeffs, commits, pieces = zo.vanishing_construct(h, t, log_n, log_ext, w_inv, divisor,
                                                     np.concatenate([zeta2, zeta]), g, q, xn)
    assert commits.shape == (q, 8)
    lhs = zo.g1_to_affine(zo.msm(g, pieces))
    acc = None
    for i in range(q):
        term = zo.g1_mul(commits[i:i + 1], zo.fr_pow(xn, i))
        acc = term if acc is None else zo.g1_add(acc, term)
    assert (lhs == zo.g1_to_affine(acc)).all()
    # and each commitment is the MSM of its own piece
    n = 1 << log_n
    assert (commits[1:2] == zo.g1_to_affine(zo.msm_naive(g, np.ascontiguousarray(coeffs[n:2 * n])))).all()
