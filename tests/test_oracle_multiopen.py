"""CPU checks pinning the oracle's multi-open restatements before the GPU is
compared with them: poly_eval (cuda/bn254.cu:238-256, :1608-1633), the
division by (X - z) of src/multiopen.rs:145-159 / :543-553, and the Shplonk
kernels shplonk_h_x_merge / shplonk_h_x_div_points (cuda/bn254.cu:1133-1170).

Each is checked against what fixes it independently of the restatement:
Horner over python integers; the DFT (evaluation at w^j is the j-th NTT
output); (X - z) q(X) + a(z) = a(X) for the division; the products written
out over python integers for the Shplonk kernels, and a round trip for the
division by the points.
"""
import numpy as np
import pytest

R = 21888242871839275222246405745257275088548364400416034343698204186575808495617


def _ints(zo, a):
    return zo.from_mont(a)


def _horner(coeffs, x):
    acc = 0
    for c in reversed(coeffs):
        acc = (acc * x + c) % R
    return acc


@pytest.mark.parametrize("log_n", [1, 2, 3, 4, 5, 8, 9, 13])
def test_poly_eval_is_horner(zo, log_n):
    n = 1 << log_n
    p = zo.rand_fr(n, 10 + log_n)
    x = zo.rand_fr(1, 20 + log_n)
    got = zo.poly_eval(p, zo.x_table(x, log_n))
    assert _ints(zo, got)[0] == _horner(_ints(zo, p), _ints(zo, x)[0])


def test_poly_eval_special_points(zo):
    log_n = 6
    n = 1 << log_n
    p = zo.rand_fr(n, 3)
    pi = _ints(zo, p)
    assert _ints(zo, zo.poly_eval(p, zo.x_table(zo.fr_zeros(1), log_n)))[0] == pi[0]
    one = zo.to_mont([1])
    assert _ints(zo, zo.poly_eval(p, zo.x_table(one, log_n)))[0] == sum(pi) % R
    # all-ones coefficients: (x^n - 1) / (x - 1)
    x = zo.rand_fr(1, 4)
    xi = _ints(zo, x)[0]
    ones = np.tile(one, (n, 1))
    exp = (pow(xi, n, R) - 1) * pow(xi - 1, -1, R) % R
    assert _ints(zo, zo.poly_eval(ones, zo.x_table(x, log_n)))[0] == exp


def test_poly_eval_at_roots_is_the_dft(zo):
    log_n = 7
    n = 1 << log_n
    p = zo.rand_fr(n, 5)
    w = zo.root_of_unity(log_n)
    dft = zo.ntt(p, log_n, w)
    wi = _ints(zo, w)[0]
    for j in (0, 1, 5, n - 1):
        xj = zo.to_mont([pow(wi, j, R)])
        assert (zo.poly_eval(p, zo.x_table(xj, log_n)) == dft[j:j + 1]).all()


@pytest.mark.parametrize("n", [1, 2, 3, 5, 64, 100])
def test_kate_divide_identity(zo, n):
    a = zo.rand_fr(n, 30 + n)
    z = zo.rand_fr(1, 40 + n)
    q = zo.kate_divide(a, z)
    ai, qi, zi = _ints(zo, a), _ints(zo, q), _ints(zo, z)[0]
    if n == 1:
        assert qi == ai  # the loop leaves one entry as it is
        return
    assert qi[n - 1] == 0
    # (X - z) q(X) + a(z) = a(X), coefficient by coefficient
    az = _horner(ai, zi)
    prod = [0] * n
    for i in range(n - 1):
        prod[i + 1] = (prod[i + 1] + qi[i]) % R
        prod[i] = (prod[i] - zi * qi[i]) % R
    prod[0] = (prod[0] + az) % R
    assert prod == ai


@pytest.mark.parametrize("diff_n", [0, 1, 3])
def test_h_x_merge_products(zo, diff_n):
    n = 40
    res, values, omegas = zo.rand_fr(n, 1), zo.rand_fr(n, 2), zo.rand_fr(n, 3)
    v = zo.rand_fr(1, 4)
    diff = zo.rand_fr(diff_n, 5) if diff_n else zo.fr_zeros(0)
    got = _ints(zo, zo.shplonk_h_x_merge(res, v, values, omegas, diff))
    ri, vi, wi, oi, di = _ints(zo, res), _ints(zo, v)[0], _ints(zo, values), _ints(zo, omegas), (
        _ints(zo, diff) if diff_n else [])
    for i in range(n):
        t = wi[i]
        for d in di:
            t = t * (oi[i] - d) % R
        assert got[i] == (ri[i] * vi + t) % R


@pytest.mark.parametrize("points_n", [1, 2, 4])
def test_h_x_div_points_round_trip(zo, points_n):
    log_n = 6
    n = 1 << log_n
    values = zo.rand_fr(n, 7)
    omegas = zo.expand_omega(n, zo.root_of_unity(log_n))
    points = zo.rand_fr(points_n, 8)
    got = _ints(zo, zo.shplonk_h_x_div_points(values, omegas, points))
    vi, oi, pi = _ints(zo, values), _ints(zo, omegas), _ints(zo, points)
    for i in range(n):
        den = 1
        for p in pi:
            den = den * (oi[i] - p) % R
        assert got[i] * den % R == vi[i]


def test_h_x_div_points_zero_denominator(zo):
    log_n = 4
    n = 1 << log_n
    values = zo.rand_fr(n, 9)
    omegas = zo.expand_omega(n, zo.root_of_unity(log_n))
    points = np.concatenate([zo.rand_fr(1, 10), omegas[5:6]])
    got = zo.shplonk_h_x_div_points(values, omegas, points)
    assert not got[5].any()
    assert got[4].any() and got[6].any()
