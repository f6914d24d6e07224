"""GPU parity: batch inversion, logup grand sums, permutation / shuffle grand products.

Every output is compared bit-for-bit with the CPU oracle on seeded inputs.
Closed-form checks from the reference: logup_sum_input_inv accumulates
exactly 1/(x + beta) per call (src/profile.rs:109-172); logup z satisfies
the lookup_z_verifiy invariants (src/lookup.rs:175-233); a valid permutation
or shuffle telescopes to a full product of 1.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

R = 21888242871839275222246405745257275088548364400416034343698204186575808495617


def _dv():
    from zkprover import device as dv
    from zkprover._lib import call
    return dv, call


def _arena():
    from zkprover.device import Arena
    return Arena()


@pytest.mark.parametrize("n", [1, 7, 255, 256, 257, 512, 513, 4095, 4096, 4097, 1 << 16, (1 << 20) + 3, 1 << 22,
                               1 << 24, (1 << 24) + 5])
def test_batch_invert(gpu, zo, n):
    """tiles of 4096 (partial last tiles, one to sixteen tile totals per
    thread of the top workgroup at 2^24) and, past 2^24, the multi-level
    hierarchy; zeros stay zero"""
    dv, call = _dv()
    A = _arena()
    x = zo.rand_fr(n, 50 + n)
    if n > 10:
        x[3] = 0
        x[n - 1] = 0
    d = dv.to_dev(x)
    tmp = dv.empty(max(n, 1))
    call("zk_batch_invert", dv.ptr(d), dv.ptr(tmp), n, dv.stream_handle())
    got = dv.from_dev(d)
    assert (got == zo.batch_invert(x)).all()
    if n >= 1 << 16:
        # x * x^-1 = 1 on every nonzero entry, zeros stay zero (closed form)
        nz = np.ones(n, dtype=bool)
        nz[[3, n - 1]] = False
        one = zo.to_mont([1])[0]
        prod = zo.field_op(x, 0, None, got, 0, None, n, 1)
        assert (prod[nz] == one).all() and (got[~nz] == 0).all()


def test_batch_invert_zero_tile(gpu, zo):
    """a tile of 4096 zeros between two tiles of values: its product is the
    empty product, its zeros stay zero, the other tiles are unaffected"""
    dv, call = _dv()
    n = 3 * 4096 + 100
    x = zo.rand_fr(n, 77)
    x[4096:8192] = 0
    d = dv.to_dev(x)
    tmp = dv.empty(n)
    call("zk_batch_invert", dv.ptr(d), dv.ptr(tmp), n, dv.stream_handle())
    got = dv.from_dev(d)
    assert (got == zo.batch_invert(x)).all()
    assert (got[4096:8192] == 0).all()


@pytest.mark.parametrize("log_n", [10, 16, 20])
def test_logup_sum_input_inv(gpu, zo, log_n):
    """the reference's KAT: after call i (init = i) the sum grows by exactly
    1/(x_j + beta) (src/profile.rs:140-170)"""
    dv, call = _dv()
    A = _arena()
    n = 1 << log_n
    beta = zo.rand_fr(1, 9)
    s = dv.empty(n)
    tmp = dv.empty(n)
    db = dv.to_dev(beta)
    exp = zo.fr_zeros(n)
    for it in range(3):
        x = zo.rand_fr(n, 100 * log_n + it)
        call("logup_sum_input_inv", dv.ptr(s), A.put(x), dv.ptr(tmp), dv.ptr(db), it, n,
             dv.stream_handle())
        exp = zo.logup_sum_input_inv(exp, x, beta, it)
        got = dv.from_dev(s)
        assert (got == exp).all(), it
    # closed form on a prefix: 1/(x + beta) summed over the three calls
    xs = [zo.rand_fr(n, 100 * log_n + it)[:64] for it in range(3)]
    bi = zo.from_mont(beta)[0]
    for j in range(64):
        v = sum(pow((zo.from_mont(x[j:j + 1])[0] + bi) % R, -1, R) for x in xs) % R
        assert zo.from_mont(got[j:j + 1])[0] == v


@pytest.mark.parametrize("log_n", [12, 16, 20])
def test_logup_z_compose_and_pure(gpu, zo, log_n):
    """set 0 via eval_logup_z_compose, set 1 via eval_logup_z_pure, chained
    through last_z (src/lib.rs:640-703)"""
    dv, call = _dv()
    A = _arena()
    n = 1 << log_n
    u = n - 7
    beta = zo.rand_fr(1, 3)
    inputs0 = [zo.rand_fr(n, 10 + k) for k in range(2)]
    inputs1 = [zo.rand_fr(n, 20 + k) for k in range(3)]
    table = zo.rand_fr(n, 30)
    m = zo.to_mont([int(v) for v in np.random.default_rng(1).integers(0, 5, n)])
    db = dv.to_dev(beta)
    tmp = dv.empty(n)
    last = dv.zeros(1)
    # set 0
    s = dv.empty(n)
    exp_s = zo.fr_zeros(n)
    for k, x in enumerate(inputs0):
        call("logup_sum_input_inv", dv.ptr(s), A.put(x), dv.ptr(tmp), dv.ptr(db), k, n,
             dv.stream_handle())
        exp_s = zo.logup_sum_input_inv(exp_s, x, beta, k)
    z0 = dv.empty(n)
    call("eval_logup_z_compose", dv.ptr(z0), dv.ptr(s), A.put(table), A.put(m), dv.ptr(db),
         dv.ptr(last), u, n, dv.stream_handle())
    exp_z0, exp_last = zo.eval_logup_z_compose(exp_s, table, m, beta, zo.fr_zeros(1), u)
    assert (dv.from_dev(z0) == exp_z0).all()
    assert (dv.from_dev(last) == exp_last).all()
    # set 1
    s1 = dv.empty(n)
    exp_s1 = zo.fr_zeros(n)
    for k, x in enumerate(inputs1):
        call("logup_sum_input_inv", dv.ptr(s1), A.put(x), dv.ptr(tmp), dv.ptr(db), k, n,
             dv.stream_handle())
        exp_s1 = zo.logup_sum_input_inv(exp_s1, x, beta, k)
    scratch_in, scratch_t = dv.empty(n), dv.empty(n)
    call("eval_logup_z_pure", dv.ptr(s1), dv.ptr(scratch_in), dv.ptr(scratch_t), dv.ptr(last), u, n,
         dv.stream_handle())
    exp_z1, exp_last1 = zo.eval_logup_z_pure(exp_s1, exp_last, u)
    got_z1 = dv.from_dev(s1)
    assert (got_z1 == exp_z1).all()
    assert (dv.from_dev(last) == exp_last1).all()
    # lookup_z_verifiy invariants (src/lookup.rs:175-233) on the GPU output
    z0h = dv.from_dev(z0)
    assert (z0h[0] == 0).all()
    assert (got_z1[0] == z0h[u]).all()


def _perm_case(zo, n, ncols, seed):
    vals = [zo.rand_fr(n, seed + j) for j in range(ncols)]
    sig = [zo.rand_fr(n, seed + 100 + j) for j in range(ncols)]
    return vals, sig


@pytest.mark.parametrize("log_n,ncols", [(10, 1), (12, 3), (16, 5), (20, 4)])
def test_permutation_z_matches_oracle(gpu, zo, log_n, ncols):
    dv, call = _dv()
    A = _arena()
    n = 1 << log_n
    beta, gamma = zo.rand_fr(1, 1), zo.rand_fr(1, 2)
    delta = zo.to_mont([pow(7, 1 << 28, R)])
    omega = zo.root_of_unity(log_n)
    u = n - 6
    zs_exp, zs_dev = [], []
    for s in range(3):
        vals, sig = _perm_case(zo, n, ncols, 1000 * s + log_n)
        dstart = zo.fr_pow(delta, s * ncols)
        exp = zo.permutation_z_set(vals, sig, beta, gamma, dstart, delta, omega)
        dvals, dsig = [dv.to_dev(v) for v in vals], [dv.to_dev(v) for v in sig]
        z = dv.empty(n)
        tmp = dv.empty(n)
        call("zk_permutation_z_set", dv.ptr(z), A.ptrs(dvals), A.ptrs(dsig), ncols,
             A.put(beta), A.put(gamma), A.put(dstart), A.put(delta),
             A.put(omega), dv.ptr(tmp), n, dv.stream_handle())
        assert (dv.from_dev(z) == exp).all(), s
        zs_exp.append(exp)
        zs_dev.append(z)
    tails = dv.empty(3)
    call("zk_permutation_chain", A.ptrs(zs_dev), 3, u, n, dv.ptr(tails), dv.stream_handle())
    chained = zo.permutation_chain(zs_exp, u)
    for s in range(3):
        assert (dv.from_dev(zs_dev[s]) == chained[s]).all(), s


def test_permutation_valid_copy_telescopes(gpu, zo):
    """a genuine copy-constraint permutation: the full product over all rows is 1"""
    dv, call = _dv()
    A = _arena()
    log_n, ncols = 12, 3
    n = 1 << log_n
    rng = np.random.default_rng(7)
    omega = zo.root_of_unity(log_n)
    delta = zo.to_mont([pow(7, 1 << 28, R)])
    w = zo.from_mont(omega)[0]
    d = zo.from_mont(delta)[0]
    # cell (j, i) has identity tag delta^j w^i; sigma permutes the tags within cycles of equal values
    cells = [(j, i) for j in range(ncols) for i in range(n)]
    perm = rng.permutation(len(cells))
    values = {}
    sigma = {}
    cyc = 4
    for c0 in range(0, len(perm), cyc):
        group = [cells[k] for k in perm[c0:c0 + cyc]]
        v = int(rng.integers(1, 1 << 60))
        for a, b in zip(group, group[1:] + group[:1]):
            values[a] = v
            sigma[a] = pow(d, b[0], R) * pow(w, b[1], R) % R
    vals = [zo.to_mont([values[(j, i)] for i in range(n)]) for j in range(ncols)]
    sig = [zo.to_mont([sigma[(j, i)] for i in range(n)]) for j in range(ncols)]
    beta, gamma = zo.rand_fr(1, 11), zo.rand_fr(1, 12)
    one = zo.fr_from_u64(1)
    z = dv.empty(n)
    tmp = dv.empty(n)
    call("zk_permutation_z_set", dv.ptr(z), A.put_all(vals),
         A.put_all(sig), ncols, A.put(beta), A.put(gamma),
         A.put(one), A.put(delta), A.put(omega), dv.ptr(tmp), n, dv.stream_handle())
    zh = dv.from_dev(z)
    assert zo.from_mont(zh[:1])[0] == 1
    assert (zh == zo.permutation_z_set(vals, sig, beta, gamma, one, delta, omega)).all()
    # the copy constraints hold, so the last row's ratio closes the product to 1
    bi, gi = zo.from_mont(beta)[0], zo.from_mont(gamma)[0]
    num = den = 1
    for j in range(ncols):
        v = values[(j, n - 1)]
        num = num * (v + bi * pow(d, j, R) * pow(w, n - 1, R) + gi) % R
        den = den * (v + bi * sigma[(j, n - 1)] + gi) % R
    assert zo.from_mont(zh[n - 1:])[0] * num * pow(den, -1, R) % R == 1


@pytest.mark.parametrize("log_n,glen", [(10, 1), (14, 2), (18, 3)])
def test_shuffle_z(gpu, zo, log_n, glen):
    dv, call = _dv()
    A = _arena()
    n = 1 << log_n
    beta = zo.rand_fr(1, 5)
    betas = np.concatenate([zo.fr_pow(beta, j + 1) for j in range(glen)])
    inputs = [zo.rand_fr(n, 40 + j) for j in range(glen)]
    perm = np.random.default_rng(3).permutation(n)
    shuffles = [np.ascontiguousarray(x[perm]) for x in inputs]
    z = dv.empty(n)
    tmp = dv.empty(n)
    call("zk_shuffle_z", dv.ptr(z), A.put_all(inputs),
         A.put_all(shuffles), glen, A.put(betas), dv.ptr(tmp), n,
         dv.stream_handle())
    zh = dv.from_dev(z)
    assert (zh == zo.shuffle_z(inputs, shuffles, beta)).all()
    # a true shuffle: the product over all n rows is 1, i.e. z[n-1] * ratio[n-1] = 1
    num = zo.fr_from_u64(1)
    den = zo.fr_from_u64(1)
    for j in range(glen):
        num = zo.fr_mul(num, zo.fr_add(inputs[j][n - 1:], betas[j:j + 1]))
        den = zo.fr_mul(den, zo.fr_add(shuffles[j][n - 1:], betas[j:j + 1]))
    closing = zo.fr_mul(zh[n - 1:], zo.fr_mul(num, zo.fr_inv(den)))
    assert zo.from_mont(closing)[0] == 1
