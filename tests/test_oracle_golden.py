"""CPU: pins the C oracle (oracle/zkoracle.c) -- no GPU.

What fixes each oracle function independently of its own code:
  * field ops, Montgomery form, inversion: golden vectors from the Python
    big-integer model (tests/golden/golden.json) and Python's own big ints;
  * G1 arithmetic, naive and Pippenger MSM: golden k*G / P+Q / 2P and small
    MSMs (with repeated and opposite points), plus the reference's test_msm
    closed form (src/cuda/msm.rs:477-511);
  * NTT / coset NTT: the DFT definition (golden DFT, the O(n^2) sum),
    NTT o iNTT = id and the convolution theorem;
  * batch inversion: golden vector with zeros, and x * x^-1 = 1;
  * logup z: golden vector and the lookup_z_verifiy invariants
    (src/lookup.rs:175-233);
  * permutation / shuffle grand products: golden vectors and telescoping of a
    genuine permutation / shuffle to 1;
  * eval_h per-row formulas: restatements only -- "parity unpinned" beyond the
    reference's formulas (cuda/bn254.cu:542-1213); see test_evalh_* below for
    algebraic cross-checks that pin two of them partially.
Each check is chosen so that a dropped term, a wrong sign or a swapped operand
in the oracle fails it.
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))

import bigint_model as M  # noqa: E402


def I(h):
    return int(h, 16)


def fr(zo, hexes):
    return zo.to_mont([I(x) for x in hexes])


def fr1(zo, h):
    return zo.to_mont([I(h)])


def aff(zo, p):
    if p is None:
        return np.zeros((1, 8), dtype=np.uint64)
    return np.concatenate([zo.fp_to_mont_int(I(p[0])), zo.fp_to_mont_int(I(p[1]))], axis=1)


def test_field_golden(zo):
    for row in GOLD["fr_ops"]:
        a, b = fr1(zo, row["a"]), fr1(zo, row["b"])
        assert zo.from_mont(zo.fr_add(a, b))[0] == I(row["add"])
        assert zo.from_mont(zo.fr_sub(a, b))[0] == I(row["sub"])
        assert zo.from_mont(zo.fr_mul(a, b))[0] == I(row["mul"])
        assert zo.from_mont(zo.fr_inv(a))[0] == I(row["inv_a"])
        assert zo.limbs_to_int(a[0]) == I(row["mont_a"])
    for row in GOLD["fp_ops"]:
        a, b = zo.fp_to_mont_int(I(row["a"])), zo.fp_to_mont_int(I(row["b"]))
        assert zo.fp_from_mont_int(zo.fp_mul(a, b)) == I(row["mul"])
        assert zo.limbs_to_int(a[0]) == I(row["mont_a"])


def test_g1_golden(zo):
    g = zo.g1_generator()
    for row in GOLD["g1_mul"]:
        got = zo.g1_to_affine(zo.g1_mul(g, fr1(zo, row["k"])))
        assert (got == aff(zo, row["kG"])).all()
        assert zo.g1a_is_on_curve(got)
    ga = GOLD["g1_add"]
    p, q = aff(zo, ga["p"]), aff(zo, ga["q"])
    pj = np.concatenate([p, zo.fp_to_mont_int(1)], axis=1)
    assert (zo.g1_to_affine(zo.g1_add_mixed(pj, q)) == aff(zo, ga["p_plus_q"])).all()
    assert (zo.g1_to_affine(zo.g1_dbl(pj)) == aff(zo, ga["two_p"])).all()
    assert (zo.g1_to_affine(zo.g1_add(pj, pj)) == aff(zo, ga["two_p"])).all()


@pytest.mark.parametrize("name", ["msm_1", "msm_4", "msm_9"])
def test_msm_golden(zo, name):
    case = GOLD[name]
    bases = np.concatenate([aff(zo, b) for b in case["bases"]])
    sc = fr(zo, case["scalars"])
    exp = aff(zo, case["result"])
    assert (zo.g1_to_affine(zo.msm_naive(bases, sc)) == exp).all()
    for c in (0, 2, 3, 7):
        assert (zo.g1_to_affine(zo.msm(bases, sc, c)) == exp).all()


def test_msm_reference_kat_golden(zo):
    case = GOLD["msm_kat_512"]
    cand = fr(zo, case["candidates"])
    g = zo.g1_generator()
    cpts = zo.g1_to_affine(np.concatenate([zo.g1_mul(g, cand[j:j + 1]) for j in range(256)]))
    pts = np.ascontiguousarray(np.tile(cpts, (2, 1)))
    sc = fr(zo, case["scalars"])
    exp = aff(zo, case["result"])
    assert (zo.g1_to_affine(zo.msm(pts, sc)) == exp).all()
    assert (zo.g1_to_affine(zo.g1_mul(g, zo.fr_inner_cycled(sc, cand))) == exp).all()


def test_msm_linearity(zo):
    n = 300
    pts = zo.gen_bases_random(n, 5)
    a, b = zo.rand_fr(n, 6), zo.rand_fr(n, 7)
    s = np.concatenate([zo.fr_add(a[i:i + 1], b[i:i + 1]) for i in range(n)])
    lhs = zo.g1_to_affine(zo.msm(pts, s))
    rhs = zo.g1_to_affine(zo.g1_add(zo.msm(pts, a), zo.msm(pts, b)))
    assert (lhs == rhs).all()


def test_bases_generators_on_curve(zo):
    for gen in (zo.gen_bases_random, zo.gen_bases_ap):
        pts = gen(64, 9)
        assert all(zo.g1a_is_on_curve(pts[i:i + 1]) for i in range(64))
        assert len({tuple(r) for r in pts}) == 64


def test_dft_golden(zo):
    case = GOLD["dft_16"]
    w = fr1(zo, case["omega"])
    x = fr(zo, case["input"])
    exp = [I(v) for v in case["output"]]
    assert zo.from_mont(zo.ntt(x, 4, w)) == exp
    assert zo.from_mont(zo.dft_naive(x, w)) == exp
    assert (zo.root_of_unity(4) == w).all()
    # the reference's s_i = s_{i-1} + i pattern generator
    s0 = zo.from_mont(zo.rand_fr(1, 3))[0]
    assert zo.from_mont(zo.arith_pattern(16, 3)) == [(s0 + i * (i + 1) // 2) % M.R for i in range(16)]


def test_coset_golden(zo):
    case = GOLD["coset_8_to_32"]
    zeta = I(case["zeta"])
    cp = zo.to_mont([zeta, zeta * zeta % M.R])
    host = np.zeros((32, 4), dtype=np.uint64)
    host[:8] = fr(zo, case["coeffs"])
    prep = zo.extended_prepare(host, cp, 8, 32, 0)
    got = zo.ntt(prep, 5, fr1(zo, case["omega"]))
    assert zo.from_mont(got) == [I(v) for v in case["output"]]


@pytest.mark.parametrize("log_n", [6, 10, 14])
def test_ntt_roundtrip_and_convolution(zo, log_n):
    n = 1 << log_n
    w = zo.root_of_unity(log_n)
    winv = zo.fr_inv(w)
    ninv = zo.fr_inv(zo.fr_from_u64(n))
    x = zo.rand_fr(n, 40 + log_n)
    assert (zo.ntt(zo.ntt(x, log_n, w), log_n, winv, ninv) == x).all()
    # convolution theorem on polynomials of degree < n/2
    h = n // 2
    a = zo.rand_fr(n, 1)
    b = zo.rand_fr(n, 2)
    a[h:] = 0
    b[h:] = 0
    A, B = zo.ntt(a, log_n, w), zo.ntt(b, log_n, w)
    C = np.concatenate([zo.fr_mul(A[i:i + 1], B[i:i + 1]) for i in range(n)])
    c = zo.ntt(C, log_n, winv, ninv)
    if log_n <= 10:
        ai, bi = zo.from_mont(a), zo.from_mont(b)
        conv = [0] * n
        for i in range(h):
            for j in range(h):
                conv[i + j] = (conv[i + j] + ai[i] * bi[j]) % M.R
        assert zo.from_mont(c) == conv
    assert zo.from_mont(c[n - 1:])[0] == 0  # degree <= n - 2


def test_batch_invert_golden(zo):
    case = GOLD["batch_invert"]
    x = fr(zo, case["input"])
    assert zo.from_mont(zo.batch_invert(x)) == [I(v) for v in case["output"]]


def test_logup_golden_and_invariants(zo):
    case = GOLD["logup_z"]
    n, u = 16, case["u"]
    beta = fr1(zo, case["beta"])
    s = zo.fr_zeros(n)
    for k, f in enumerate(case["inputs0"]):
        s = zo.logup_sum_input_inv(s, fr(zo, f), beta, k)
    assert zo.from_mont(s) == [I(v) for v in case["sum0"]]
    z0, last0 = zo.eval_logup_z_compose(s, fr(zo, case["table"]), zo.to_mont(case["m"]), beta,
                                        zo.fr_zeros(1), u)
    assert zo.from_mont(z0) == [I(v) for v in case["z0"]]
    assert zo.from_mont(last0)[0] == I(case["last0"])
    s1 = zo.fr_zeros(n)
    for k, f in enumerate(case["inputs1"]):
        s1 = zo.logup_sum_input_inv(s1, fr(zo, f), beta, k)
    z1, last1 = zo.eval_logup_z_pure(s1, last0, u)
    assert zo.from_mont(z1) == [I(v) for v in case["z1"]]
    assert zo.from_mont(last1)[0] == I(case["last1"])
    # lookup_z_verifiy: z0[0] = 0, z1[0] = z0[u], z[i+1] = z[i] + S_i - m_i/(t_i+beta)
    zi0 = zo.from_mont(z0)
    assert zi0[0] == 0 and zo.from_mont(z1)[0] == zi0[u]
    b = I(case["beta"])
    for i in range(n - 1):
        step = (I(case["sum0"][i]) - case["m"][i] * M.inv(I(case["table"][i]) + b, M.R)) % M.R
        assert zi0[i + 1] == (zi0[i] + step) % M.R


def test_permutation_golden(zo):
    case = GOLD["permutation_z"]
    z = zo.permutation_z_set([fr(zo, c) for c in case["values"]], [fr(zo, c) for c in case["sigmas"]],
                             fr1(zo, case["beta"]), fr1(zo, case["gamma"]), fr1(zo, case["delta_start"]),
                             fr1(zo, case["delta"]), fr1(zo, case["omega"]))
    assert zo.from_mont(z) == [I(v) for v in case["z"]]


def test_permutation_chain_semantics(zo):
    """z_s[0..=u] *= prod_{s'<s} z_s'[u] (src/permutation.rs:131-147)"""
    n, u = 32, 25
    zs = [zo.rand_fr(n, 70 + s) for s in range(3)]
    for z in zs:
        z[0] = zo.fr_from_u64(1)[0]  # grand products start at 1
    ch = zo.permutation_chain(zs, u)
    ints = [zo.from_mont(z) for z in zs]
    out = [zo.from_mont(z) for z in ch]
    assert out[0] == ints[0]
    t0 = ints[0][u]
    assert out[1][: u + 1] == [v * t0 % M.R for v in ints[1][: u + 1]] and out[1][u + 1:] == ints[1][u + 1:]
    t01 = t0 * ints[1][u] % M.R
    assert out[2][: u + 1] == [v * t01 % M.R for v in ints[2][: u + 1]]
    assert out[2][0] == out[1][u]  # the chaining property itself


def test_shuffle_golden_and_telescoping(zo):
    case = GOLD["shuffle_z"]
    beta = fr1(zo, case["beta"])
    z = zo.shuffle_z([fr(zo, c) for c in case["inputs"]], [fr(zo, c) for c in case["shuffles"]], beta)
    assert zo.from_mont(z) == [I(v) for v in case["z"]]
    # a genuine shuffle telescopes to 1
    n = 64
    ins = [zo.rand_fr(n, 90 + j) for j in range(2)]
    perm = np.random.default_rng(4).permutation(n)
    shf = [np.ascontiguousarray(x[perm]) for x in ins]
    z = zo.shuffle_z(ins, shf, beta)
    zi = zo.from_mont(z)
    bi = I(case["beta"])
    last = M.shuffle_ratio_row([zo.from_mont(x[n - 1:])[0] for x in ins], [zo.from_mont(x[n - 1:])[0] for x in shf], bi)
    assert zi[n - 1] * last % M.R == 1


def test_evalh_formulas_against_model(zo):
    """per-row restatements of cuda/bn254.cu:664-1071 against a direct evaluation
    in Python ints of the same formulas (parity unpinned beyond the formulas)"""
    n, rot = 16, 4
    rng = np.random.default_rng(8)
    arrs = [zo.rand_fr(n, 200 + k) for k in range(10)]
    ints = [zo.from_mont(a) for a in arrs]
    y = zo.rand_fr(1, 300)
    yi = zo.from_mont(y)[0]
    res0 = arrs[0]
    r = ints[0]
    got = zo.from_mont(zo.permutation_eval_h_p1(res0, arrs[1], arrs[2], arrs[3], arrs[4], y))
    exp = [((r[i] * yi + (1 - ints[1][i]) * ints[3][i]) * yi + (ints[2][i] ** 2 - ints[2][i]) * ints[4][i]) % M.R
           for i in range(n)]
    assert got == exp
    got = zo.from_mont(zo.shuffle_eval_h(res0, arrs[1], arrs[2], arrs[3], arrs[4], arrs[5], arrs[6], y, rot))
    z = ints[3]
    exp = []
    for i in range(n):
        t = (r[i] * yi + ints[4][i] * (1 - z[i])) % M.R
        t = (t * yi + ints[5][i] * (z[i] * z[i] - z[i])) % M.R
        t = (t * yi + (ints[2][i] * z[(i + rot) % n] - ints[1][i] * z[i]) * ints[6][i]) % M.R
        exp.append(t)
    assert got == exp
    del rng


def test_field_op_batch_mul_sum_model(zo):
    n = 16
    cols = [zo.rand_fr(n, 10 + j) for j in range(3)]
    ci = [zo.from_mont(c) for c in cols]
    c0, c1 = zo.rand_fr(1, 20), zo.rand_fr(1, 21)
    res = zo.rand_fr(n, 22)
    got = zo.from_mont(zo.field_op_batch_mul_sum(res, [(c0, [(cols[0], 1), (cols[1], -2)]), (c1, [])], n))
    k0, k1 = zo.from_mont(c0)[0], zo.from_mont(c1)[0]
    ri = zo.from_mont(res)
    exp = [(ri[i] + k0 * ci[0][(i + 1) % n] * ci[1][(i - 2) % n] + k1) % M.R for i in range(n)]
    assert got == exp
