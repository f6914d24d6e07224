"""CPU: the oracle's eval_h constraint kernels pinned by what the mathematics
fixes, independently of the reference's code.

For each argument an honest witness is built in python integers over a
64-row domain (u = 58 usable rows, l0 / l_last / l_active as halo2 defines
them): logup grand sums (z(wX) - z(X) = sum 1/(f_j + beta) - m / (t + beta),
z starting and ending at 0), extra logup input sets, chained z sets,
permutation first / last products, shuffle grand products for one group and
for several groups with their own betas.  On such a witness every constraint
vanishes on every row, so the kernel must return exactly its accumulator
times y^k (k = the number of folds); with one cell of the witness changed
some row must differ.  Also: the product-sum helper equals
prod * sum_j 1 / phi_j, and permutation_eval_h_l adds beta p(X) + gamma as
polynomials (checked by evaluation at a random point).  field_sum is the
rotation theorem: scaling coefficient i by w^(r i) shifts the evaluations by
r.  lookup_eval_h (the permuted-column lookup) vanishes on a permuted pair
built as halo2 builds it (sorted inputs, each value's table row at its first
occurrence) and fails when two table rows are exchanged.  On arbitrary values
shuffle_eval_h_v2 over g groups equals shuffle_eval_h on the group products,
which ties v2's fold order (the one the prover core's proofs check) to
shuffle_eval_h and to several groups.
"""
import numpy as np
import pytest

R = 21888242871839275222246405745257275088548364400416034343698204186575808495617
N = 64
U = N - 6  # usable rows: [0, U); l_last at U


def inv(a):
    return pow(a % R, -1, R)


@pytest.fixture(scope="module")
def env(zo):
    rng = np.random.default_rng(2718)

    def rand(k):
        return [int.from_bytes(rng.bytes(32), "little") % R for _ in range(k)]

    def m(vals):
        return zo.to_mont([v % R for v in vals])

    def ints(arr):
        return [int(v) for v in zo.from_mont(arr)]

    l0 = [1] + [0] * (N - 1)
    ll = [0] * N
    ll[U] = 1
    la = [1] * U + [0] * (N - U)
    return zo, rand, m, ints, l0, ll, la


def _expect_fold(ints_out, res0, y, k):
    return ints_out == [r * pow(y, k, R) % R for r in res0]


def _logup_witness(rand, nin, seed_shift=0):
    """inputs drawn from a table of 16 values; m counts them per matching
    table row (the largest usable row holding the value); z from the sum"""
    table_vals = [v for v in rand(16)]
    t = [table_vals[i % 16] for i in range(N)]
    rng = np.random.default_rng(99 + seed_shift)
    f = [[table_vals[int(rng.integers(0, 16))] for _ in range(N)] for _ in range(nin)]
    last_row = {}
    for i in range(U):
        last_row[t[i]] = i
    mult = [0] * N
    for j in range(nin):
        for i in range(U):
            mult[last_row[f[j][i]]] += 1
    beta = rand(1)[0]
    z = [0] * N
    for i in range(U):
        s = sum(inv(f[j][i] + beta) for j in range(nin)) - mult[i] * inv(t[i] + beta)
        z[i + 1] = (z[i] + s) % R
    for i in range(U + 1, N):
        z[i] = rand(1)[0]  # blinding rows
    return f, t, mult, beta, z


def _prod_sum(phis):
    prod = [1] * N
    psum = [0] * N
    for i in range(N):
        for p in phis:
            prod[i] = prod[i] * p[i] % R
        for j in range(len(phis)):
            u = 1
            for k, p in enumerate(phis):
                if k != j:
                    u = u * p[i] % R
            psum[i] = (psum[i] + u) % R
    return prod, psum


@pytest.mark.parametrize("nin", [1, 2, 3])
def test_logup_vanishes_on_honest_witness(env, nin):
    zo, rand, m, ints, l0, ll, la = env
    f, t, mult, beta, z = _logup_witness(rand, nin, nin)
    assert z[U] == 0, "a valid logup sums to zero over the usable rows"
    phis = [[(x + beta) % R for x in fj] for fj in f]
    prod, psum = _prod_sum(phis)
    tb = [(x + beta) % R for x in t]
    y = rand(1)[0]
    res0 = rand(N)
    args = lambda zz, mm: (m(res0), m(prod), m(psum), m(tb), m(mm), m(zz), m(zz), m(l0), m(ll), m(la), m([y]), 1)
    out = ints(zo.logup_eval_h(*args(z, mult)))
    assert _expect_fold(out, res0, y, 3)
    # one multiplicity off by one: the sum no longer closes
    bad = list(mult)
    bad[3] += 1
    zb = [0] * N
    for i in range(U):
        s = sum(inv(f[j][i] + beta) for j in range(nin)) - bad[i] * inv(t[i] + beta)
        zb[i + 1] = (zb[i] + s) % R
    zb[U + 1:] = z[U + 1:]
    assert not _expect_fold(ints(zo.logup_eval_h(*args(zb, bad))), res0, y, 3)
    # one input changed, z kept: the row's step no longer matches
    phis[0][7] = (phis[0][7] + 1) % R
    prod2, psum2 = _prod_sum(phis)
    out2 = ints(zo.logup_eval_h(m(res0), m(prod2), m(psum2), m(tb), m(mult), m(z), m(z), m(l0), m(ll), m(la),
                                m([y]), 1))
    diff = [i for i in range(N) if out2[i] != out[i]]
    assert diff == [7]


def test_logup_extra_inputs_vanish(env):
    zo, rand, m, ints, l0, ll, la = env
    phis = [rand(N) for _ in range(3)]
    prod, psum = _prod_sum(phis)
    z = [0] * N
    for i in range(U):
        z[i + 1] = (z[i] + sum(inv(p[i]) for p in phis)) % R
    y = rand(1)[0]
    res0 = rand(N)
    out = ints(zo.logup_eval_h_extra_inputs(m(res0), m(prod), m(psum), m(z), m(la), m([y]), 1))
    assert _expect_fold(out, res0, y, 1)
    z[9] = (z[9] + 1) % R
    out = ints(zo.logup_eval_h_extra_inputs(m(res0), m(prod), m(psum), m(z), m(la), m([y]), 1))
    assert [i for i in range(N) if out[i] != res0[i] * y % R] == [8, 9]


def test_inputs_product_sum_identity(env):
    zo, rand, m, ints, l0, ll, la = env
    for k in (1, 2, 4):
        sets = [rand(N) for _ in range(k)]
        prod, psum = zo.logup_eval_h_inputs_product_sum([m(s) for s in sets])
        prod, psum = ints(prod), ints(psum)
        for i in range(N):
            p = 1
            for s in sets:
                p = p * s[i] % R
            assert prod[i] == p
            assert psum[i] == p * sum(inv(s[i]) for s in sets) % R


@pytest.mark.parametrize("nset", [2, 3, 4])
def test_z_set_chain_vanishes(env, nset):
    """sets chained as z_j(w^0) = z_(j-1)(w^U): the l0-gated link terms"""
    zo, rand, m, ints, l0, ll, la = env
    sets = [rand(N) for _ in range(nset)]
    for j in range(1, nset):
        sets[j][0] = sets[j - 1][U]
    y = rand(1)[0]
    res0 = rand(N)
    out = ints(zo.permutation_eval_h_p2(m(res0), [m(s) for s in sets], m(l0), m([y]), U))
    assert _expect_fold(out, res0, y, nset - 1)
    sets[nset - 1][0] = (sets[nset - 1][0] + 1) % R
    out = ints(zo.permutation_eval_h_p2(m(res0), [m(s) for s in sets], m(l0), m([y]), U))
    assert [i for i in range(N) if out[i] != res0[i] * pow(y, nset - 1, R) % R] == [0]


def test_permutation_first_last(env):
    zo, rand, m, ints, l0, ll, la = env
    first, last = rand(N), rand(N)
    first[0] = 1
    for bit in (0, 1):
        last[U] = bit
        y = rand(1)[0]
        res0 = rand(N)
        out = ints(zo.permutation_eval_h_p1(m(res0), m(first), m(last), m(l0), m(ll), m([y])))
        assert _expect_fold(out, res0, y, 2)
    last[U] = 2
    out = ints(zo.permutation_eval_h_p1(m(res0), m(first), m(last), m(l0), m(ll), m([y])))
    assert [i for i in range(N) if out[i] != res0[i] * y * y % R] == [U]


def test_permutation_l_adds_polynomials(env):
    zo, rand, m, ints, l0, ll, la = env
    res0, p = rand(N), rand(N)
    beta, gamma, x = rand(3)
    out = ints(zo.permutation_eval_h_l(m(res0), m([beta]), m([gamma]), m(p)))
    ev = lambda c: sum(ci * pow(x, i, R) for i, ci in enumerate(c)) % R
    assert ev(out) == (ev(res0) + beta * ev(p) + gamma) % R


def _shuffle_z(num, den):
    z = [1] + [0] * (N - 1)
    for i in range(U):
        z[i + 1] = z[i] * num[i] % R * inv(den[i]) % R
    return z


def test_shuffle_vanishes_on_honest_witness(env):
    zo, rand, m, ints, l0, ll, la = env
    vals = rand(N)
    perm = np.random.default_rng(5).permutation(U)
    inp = list(vals)
    tab = [vals[perm[i]] if i < U else vals[i] for i in range(N)]
    z = _shuffle_z(inp, tab)
    assert z[U] == 1, "a permutation of the usable rows closes the product"
    for i in range(U + 1, N):
        z[i] = rand(1)[0]
    y = rand(1)[0]
    res0 = rand(N)
    out = ints(zo.shuffle_eval_h(m(res0), m(inp), m(tab), m(z), m(l0), m(ll), m(la), m([y]), 1))
    assert _expect_fold(out, res0, y, 3)
    tab[4] = (tab[4] + 1) % R
    out = ints(zo.shuffle_eval_h(m(res0), m(inp), m(tab), m(z), m(l0), m(ll), m(la), m([y]), 1))
    assert [i for i in range(N) if out[i] != res0[i] * pow(y, 3, R) % R] == [4]


@pytest.mark.parametrize("g", [1, 2, 3])
def test_shuffle_v2_vanishes_on_honest_witness(env, g):
    zo, rand, m, ints, l0, ll, la = env
    betas = rand(g)
    ins, tabs = [], []
    for j in range(g):
        vals = rand(N)
        perm = np.random.default_rng(40 + j).permutation(U)  # each group its own permutation
        ins.append(list(vals))
        tabs.append([vals[perm[i]] if i < U else vals[i] for i in range(N)])
    num = [1] * N
    den = [1] * N
    for i in range(N):
        for j in range(g):
            num[i] = num[i] * (ins[j][i] + betas[j]) % R
            den[i] = den[i] * (tabs[j][i] + betas[j]) % R
    z = _shuffle_z(num, den)
    assert z[U] == 1
    for i in range(U + 1, N):
        z[i] = rand(1)[0]
    y = rand(1)[0]
    res0 = rand(N)
    call = lambda tb: ints(zo.shuffle_eval_h_v2(m(res0), [m(x) for x in ins], [m(x) for x in tb], m(betas), m(z),
                                                m(l0), m(ll), m(la), m([y]), 1))
    assert _expect_fold(call(tabs), res0, y, 3)
    tabs[g - 1][11] = (tabs[g - 1][11] + 1) % R
    out = call(tabs)
    assert [i for i in range(N) if out[i] != res0[i] * pow(y, 3, R) % R] == [11]


def test_field_sum_is_rotation_of_coefficients(env):
    """multiplying coefficient i by w^(r i) evaluates p at w^r X: the NTT of
    field_sum's output is sum_j c_j NTT(v_j) shifted by r_j"""
    zo, rand, m, ints, l0, ll, la = env
    log_n = 6
    w = zo.root_of_unity(log_n)
    wi = zo.from_mont(w)[0]
    omegas = m([pow(wi, i, R) for i in range(N)])
    vs = [m(rand(N)) for _ in range(4)]
    rots = [0, 1, -1, 5]
    cs = [None, m(rand(1)), m(rand(1)), None]
    out = zo.field_sum(vs, cs, rots, omegas)
    got = ints(zo.ntt(out, log_n, w))
    evs = [ints(zo.ntt(v, log_n, w)) for v in vs]
    cints = [1 if c is None else ints(c)[0] for c in cs]
    for k in range(N):
        assert got[k] == sum(cints[j] * evs[j][(k + rots[j]) % N] for j in range(4)) % R, k
    assert (zo.field_sum(vs[:0] or [m([0] * N)], [None], [0], omegas) == 0).all()


def _plookup_witness(rand):
    """halo2's permuted pair: a' the sorted inputs, s' the table with each
    input value placed at the value's first row in a', the rest anywhere;
    z the product of (a + beta)(t + gamma) / ((a' + beta)(s' + gamma))"""
    tvals = rand(U)
    rng = np.random.default_rng(7)
    a = [tvals[int(rng.integers(0, 12))] for _ in range(U)]
    ap = sorted(a)
    sp = [None] * U
    rest = list(tvals)
    for i in range(U):
        if i == 0 or ap[i] != ap[i - 1]:
            sp[i] = ap[i]
            rest.remove(ap[i])
    for i in range(U):
        if sp[i] is None:
            sp[i] = rest.pop()
    beta, gamma = rand(2)
    z = [1] + [0] * (N - 1)
    for i in range(U):
        num = (a[i] + beta) * (tvals[i] + gamma) % R
        den = (ap[i] + beta) * (sp[i] + gamma) % R
        z[i + 1] = z[i] * num % R * inv(den) % R
    pad = rand(N - U)
    return a + pad, tvals + pad, ap + pad, sp + pad, z, beta, gamma


def test_lookup_eval_h_vanishes_on_permuted_pair(env):
    zo, rand, m, ints, l0, ll, la = env
    a, t, ap, sp, z, beta, gamma = _plookup_witness(rand)
    assert z[U] == 1, "a valid permuted pair closes the product"
    for i in range(U + 1, N):
        z[i] = rand(1)[0]
    inp = [(x + beta) % R for x in a]
    tab = [(x + gamma) % R for x in t]
    y = rand(1)[0]
    res0 = rand(N)
    call = lambda spv: ints(zo.lookup_eval_h(m(res0), m(inp), m(tab), m(ap), m(spv), m(z), m(l0), m(ll), m(la),
                                             m([y]), m([beta]), m([gamma]), 1))
    assert _expect_fold(call(sp), res0, y, 5)
    # s' rows 0 and 1 exchanged: a'(0) = s'(0) fails, and so does the
    # adjacency rule wherever the moved table value lands
    bad = list(sp)
    j = next(i for i in range(1, U) if ap[i] != ap[0])
    bad[0], bad[j] = bad[j], bad[0]
    out = call(bad)
    assert out[0] != res0[0] * pow(y, 5, R) % R


@pytest.mark.parametrize("g,rot", [(1, 1), (2, 4), (3, 1), (16, 4)])
def test_shuffle_v2_is_shuffle_on_group_products(env, g, rot):
    """on arbitrary (not honest) values, shuffle_eval_h_v2 over g groups equals
    shuffle_eval_h with input prod_j (a_j + beta_j) and table prod_j (s_j +
    beta_j), row by row: v2 folds the same three terms in the same order.  The
    prover core calls v2 with one group, so its fold order is the one the
    proof's quotient identity checks; this carries that to shuffle_eval_h and
    to several groups"""
    zo, rand, m, ints, l0, ll, la = env
    betas = rand(g)
    ins = [rand(N) for _ in range(g)]
    tabs = [rand(N) for _ in range(g)]
    z, res0, y = rand(N), rand(N), rand(1)[0]
    num, den = [1] * N, [1] * N
    for i in range(N):
        for j in range(g):
            num[i] = num[i] * (ins[j][i] + betas[j]) % R
            den[i] = den[i] * (tabs[j][i] + betas[j]) % R
    v2 = zo.shuffle_eval_h_v2(m(res0), [m(x) for x in ins], [m(x) for x in tabs], m(betas), m(z), m(l0), m(ll),
                              m(la), m([y]), rot)
    one = zo.shuffle_eval_h(m(res0), m(num), m(den), m(z), m(l0), m(ll), m(la), m([y]), rot)
    assert ints(v2) == ints(one)
    # a swapped pair of groups' betas changes the products, so some row differs
    if g > 1:
        sw = betas[1:2] + betas[:1] + betas[2:]
        v2b = zo.shuffle_eval_h_v2(m(res0), [m(x) for x in ins], [m(x) for x in tabs], m(sw), m(z), m(l0), m(ll),
                                   m(la), m([y]), rot)
        assert ints(v2b) != ints(one)
