"""GPU: the prover core's openings verify as KZG openings.

The key's SRS is a KZG setup for a seeded secret tau (g[i] = tau^i G,
g_lag[i] = L_i(tau) G), so every commitment is [p(tau)] G.  With the
trapdoor the verifier's pairing equations become equations in G1, checked
here with the oracle's curve arithmetic:
  * the two bases agree: committing coefficients against g and the same
    polynomial's values against g_lag give one point (for the fixed columns
    of the verifying key, rebuilt from the key's own SRS);
  * GWC (src/multiopen.rs:68-177): for each opening point z, with the batch
    C = sum v^(len-1-i) C_i and E = sum v^(len-1-i) e_i over the queries at z,
    (tau - z) W_z = C - E G;
  * Shplonk (src/multiopen.rs:283-561): with V_s the y-combined commitments
    of each point set, r_s the interpolation of its evaluations and z_s the
    product of (u - p) over the points outside the set,
    (tau - u) [L] = (sum_s v^(k-1-s) z_s (V_s - r_s(u) G) - Z_T(u) [h]) / z_0.
A wrong evaluation, quotient or challenge breaks the equation for all but a
negligible set of tau.  The same openings also pass the verifier's own
pairing checks, which hold only [tau]_2 = tau G2 of the setup (the pairing of
oracle/bn254_pairing.py): GWC as one product check over all points with a
random u, Shplonk as e([L], [tau]_2) = e(u [L] + F / z_0, G2); the tampered
proof fails it.  The vanishing argument is checked the verifier's way
too: h(x) (x^n - 1) equals the composed constraint rebuilt from the opened
evaluations and the challenges.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

R = 21888242871839275222246405745257275088548364400416034343698204186575808495617


def _circuit():
    from zkprover.prover import Circuit
    return Circuit(k=10, advice=30, fixed=2, instance=1, lookups=3, table_log=7, perm_cols=6, shuffles=2,
                   shuffle_log=7, blinding=5)


@pytest.fixture(scope="module")
def setup(gpu):
    from zkprover.prover import Prover, ProvingKey
    pk = ProvingKey(_circuit(), seed=9)
    pr = Prover(pk)
    yield pk, pr
    pr.close()
    pk.close()


class G1:
    """curve arithmetic of the oracle on (1, 12) Jacobian arrays"""

    def __init__(self, zo):
        self.zo = zo
        self.g = zo.g1_generator()

    def mul_aff(self, aff, s):
        return self.zo.g1_mul(np.ascontiguousarray(aff, dtype=np.uint64).reshape(1, 8),
                              self.zo.to_mont([s % R]))

    def add(self, a, b):
        return self.zo.g1_add(a, b)

    def lin(self, pairs):
        """sum s_i P_i for (affine P, int s) pairs"""
        acc = None
        for p, s in pairs:
            t = self.mul_aff(p, s)
            acc = t if acc is None else self.add(acc, t)
        return acc

    def aff(self, jac):
        return self.zo.g1_to_affine(jac)


def _ints(zo, arr):
    return zo.from_mont(np.ascontiguousarray(arr, dtype=np.uint64).reshape(-1, 4))


def _layout(circ, com, vk, x, g1):
    """commitments of the opened polynomials per point set, in query order"""
    A, L, P, S = circ.advice, circ.lookups, circ.perm_cols, circ.shuffles
    Sp = (P + 1) // 2
    adv, m = com[:A], com[A:A + L]
    zp = com[A + L:A + L + Sp]
    zl = com[A + L + Sp:A + 2 * L + Sp]
    zs = com[A + 2 * L + Sp:A + 2 * L + Sp + S]
    pieces = com[A + 2 * L + Sp + S:A + 2 * L + Sp + S + 3]
    n = 1 << circ.k
    xn = pow(x, n, R)
    h = g1.aff(g1.lin([(pieces[i], pow(xn, i, R)) for i in range(3)]))
    sets = [list(adv) + [vk[0], vk[1]] + list(vk[2:2 + P]) + list(m) + [h[0]],
            list(zl) + list(zs) + [zp[Sp - 1]],
            list(zp[:Sp - 1])]
    return sets, A + 2 * L + Sp + S + 3


def _points(circ, x, zo):
    n = 1 << circ.k
    w = _ints(zo, zo.root_of_unity(circ.k))[0]
    u = n - (circ.blinding + 1)
    return [x, x * w % R, x * pow(w, u, R) % R]


def _evals_by_set(sets, ev):
    out, e = [], 0
    for s, polys in enumerate(sets):
        cur = []
        for _ in polys:
            cur.append(ev[e:e + s + 1])
            e += s + 1
        out.append(cur)
    assert e == len(ev)
    return out


def test_fixed_commitments_agree_across_bases(setup, zo):
    """the verifying key (coefficients against g) equals the Lagrange-basis
    commitment of the same columns' values: g and g_lag come from one tau"""
    pk, _ = setup
    tau = _ints(zo, pk.srs_secret())[0]
    vk = pk.fixed_commitments()
    assert len(vk) == 2 + pk.circuit.perm_cols
    assert all(zo.g1a_is_on_curve(vk[i:i + 1]) for i in range(len(vk)))
    # q is l_active: 1 on rows [0, u); its polynomial at tau is sum_{i<u} L_i(tau)
    circ = pk.circuit
    n, u = 1 << circ.k, (1 << circ.k) - (circ.blinding + 1)
    w = _ints(zo, zo.root_of_unity(circ.k))[0]
    c = (pow(tau, n, R) - 1) * pow(n, -1, R) % R
    q_tau = sum(pow(w, i, R) * pow(tau - pow(w, i, R), -1, R) for i in range(u)) * c % R
    g1 = G1(zo)
    assert (g1.aff(g1.mul_aff(g1.g, q_tau)) == vk[0:1]).all()


def _gwc_groups(pk, com, info, ev, v, zo, g1):
    """per opening point q: (W_q affine row, z_q, C_q - E_q G Jacobian) with
    the v-batched commitment C_q and evaluation E_q of the queries at z_q"""
    x = _ints(zo, np.array(info.challenges[3], dtype=np.uint64))[0]
    sets, wbase = _layout(pk.circuit, com, pk.fixed_commitments(), x, g1)
    pts = _points(pk.circuit, x, zo)
    evs = _evals_by_set(sets, ev)
    ngroups = 3 if sets[2] else (2 if sets[1] else 1)
    out = []
    for q in range(ngroups):
        pairs, evals = [], []
        for s in range(q, 3):
            for i, c in enumerate(sets[s]):
                pairs.append(c)
                evals.append(evs[s][i][q])
        k = len(pairs)
        batch = g1.lin([(pairs[i], pow(v, k - 1 - i, R)) for i in range(k)])
        e = sum(evals[i] * pow(v, k - 1 - i, R) for i in range(k)) % R
        out.append((com[wbase + q], pts[q], g1.add(batch, g1.mul_aff(g1.g, -e))))
    return out


def _gwc_pairing_accepts(groups, tau, g1, u=0x5eed):
    """the verifier's one pairing check over all points, holding only
    [tau]_2 of the setup: with powers of a random u,
    e(sum u^q W_q, [tau]_2) = e(sum u^q (z_q W_q + C_q - E_q G), G2)"""
    import bn254_pairing as bp
    left = g1.lin([(w, pow(u, q, R)) for q, (w, _, _) in enumerate(groups)])
    right = None
    for q, (w, z, rhs) in enumerate(groups):
        t = g1.mul_aff(g1.aff(g1.add(g1.mul_aff(w, z), rhs))[0], pow(u, q, R))
        right = t if right is None else g1.add(right, t)
    pt = lambda jac: None if not g1.aff(jac).any() else g1.zo.g1a_to_ints(g1.aff(jac))
    return bp.pairing_check([(bp.g2_mul(bp.G2, tau), pt(left)), (bp.g2_neg(bp.G2), pt(right))])


@pytest.mark.parametrize("seed", [41, 42])
def test_gwc_openings_verify(setup, zo, seed):
    from zkprover.prover import BATCHED
    pk, pr = setup
    com, _, info = pr.prove(seed, flags=BATCHED)
    assert info.open_ok == 1
    ev_m, ch = pr.openings()
    ev, v = _ints(zo, ev_m), _ints(zo, ch)[0]
    tau = _ints(zo, pk.srs_secret())[0]
    g1 = G1(zo)
    groups = _gwc_groups(pk, com, info, ev, v, zo, g1)
    for q, (w, z, rhs) in enumerate(groups):
        lhs = g1.mul_aff(w, tau - z)
        assert (g1.aff(lhs) == g1.aff(rhs)).all(), "GWC opening at point %d" % q
    if seed == 41:
        assert _gwc_pairing_accepts(groups, tau, g1)


def _lagrange_at(xs, ys, t):
    acc = 0
    for k, (xk, yk) in enumerate(zip(xs, ys)):
        num, den = 1, 1
        for l, xl in enumerate(xs):
            if l != k:
                num = num * (t - xl) % R
                den = den * (xk - xl) % R
        acc = (acc + yk * num * pow(den, -1, R)) % R
    return acc


def test_shplonk_opening_verifies(setup, zo):
    from zkprover.prover import BATCHED, SHPLONK
    pk, pr = setup
    com, _, info = pr.prove(43, flags=BATCHED | SHPLONK)
    assert info.open_ok == 1
    ev_m, ch = pr.openings()
    ev = _ints(zo, ev_m)
    v, y, u = _ints(zo, ch)
    tau = _ints(zo, pk.srs_secret())[0]
    x = _ints(zo, np.array(info.challenges[3], dtype=np.uint64))[0]
    g1 = G1(zo)
    sets, base = _layout(pk.circuit, com, pk.fixed_commitments(), x, g1)
    hx, lx = com[base], com[base + 1]
    pts = _points(pk.circuit, x, zo)
    evs = _evals_by_set(sets, ev)
    nsets = 3 if sets[2] else (2 if sets[1] else 1)
    zt = 1
    for q in range(nsets):
        zt = zt * (u - pts[q]) % R
    F, z0 = None, None
    for s in range(nsets):
        k = len(sets[s])
        vs = g1.lin([(sets[s][i], pow(y, k - 1 - i, R)) for i in range(k)])
        ecomb = [sum(evs[s][i][q] * pow(y, k - 1 - i, R) for i in range(k)) % R for q in range(s + 1)]
        ru = _lagrange_at(pts[:s + 1], ecomb, u)
        zs = 1
        for q in range(s + 1, nsets):
            zs = zs * (u - pts[q]) % R
        if s == 0:
            z0 = zs
        term = g1.add(vs, g1.mul_aff(g1.g, -ru))
        term = g1.mul_aff(g1.aff(term)[0], zs * pow(v, nsets - 1 - s, R))
        F = term if F is None else g1.add(F, term)
    F = g1.add(F, g1.mul_aff(hx, -zt))
    rhs = g1.mul_aff(g1.aff(F)[0], pow(z0, -1, R))
    lhs = g1.mul_aff(lx, tau - u)
    assert (g1.aff(lhs) == g1.aff(rhs)).all()
    # the verifier's pairing form: e([L], [tau]_2) = e(u [L] + F / z_0, G2)
    import bn254_pairing as bp
    right = g1.aff(g1.add(g1.mul_aff(lx, u), rhs))
    assert bp.pairing_check([(bp.g2_mul(bp.G2, tau), zo.g1a_to_ints(lx)),
                             (bp.g2_neg(bp.G2), zo.g1a_to_ints(right))])


def test_tampered_evaluation_fails_verification(setup, zo):
    """one evaluation off by one: the GWC equation at its point must fail"""
    from zkprover.prover import BATCHED, TAMPER_EVAL
    pk, pr = setup
    com, _, info = pr.prove(44, flags=BATCHED | TAMPER_EVAL)
    ev_m, ch = pr.openings()
    ev, v = _ints(zo, ev_m), _ints(zo, ch)[0]
    tau = _ints(zo, pk.srs_secret())[0]
    g1 = G1(zo)
    groups = _gwc_groups(pk, com, info, ev, v, zo, g1)
    ok = [bool((g1.aff(g1.mul_aff(w, tau - z)) == g1.aff(rhs)).all()) for w, z, rhs in groups]
    assert not all(ok)
    assert not _gwc_pairing_accepts(groups, tau, g1)


# delta = 7^(2^28) mod r, the generator of the permutation column cosets
DELTA = 0x09226b6e22c6f0ca64ec26aad4c86e715b5f898e5e963f25870e56bbe533e9a2


def _constraints_at_x(circ, evs, x, beta, gamma, y, zo):
    """the composed constraint C(x) as the prover folds it on the coset (gate;
    permutation p1, p2, one term per set; logup per lookup; shuffle per
    pair; each new term after multiplying the sum so far by y), restated
    with the opened evaluations and the verifier's own l0, l_last, l_active
    at x"""
    A, L, P, S = circ.advice, circ.lookups, circ.perm_cols, circ.shuffles
    Sp = (P + 1) // 2
    n = 1 << circ.k
    u = n - (circ.blinding + 1)
    w = _ints(zo, zo.root_of_unity(circ.k))[0]
    zh = (pow(x, n, R) - 1) % R
    ninv = pow(n, -1, R)

    def lag(i):
        wi = pow(w, i, R)
        return wi * zh % R * ninv % R * pow(x - wi, -1, R) % R

    l0, l_last = lag(0), lag(u)
    l_active = (1 - sum(lag(i) for i in range(u, n))) % R
    s0 = evs[0]
    adv = [s0[i][0] for i in range(A)]
    q, t = s0[A][0], s0[A + 1][0]
    sig = [s0[A + 2 + c][0] for c in range(P)]
    m = [s0[A + 2 + P + l][0] for l in range(L)]
    zl = [evs[1][l] for l in range(L)]
    zs = [evs[1][L + i] for i in range(S)]
    zp = [evs[2][j] for j in range(Sp - 1)] + [evs[1][L + S]]
    a, b, c, d = adv[0:4]
    h = (q * a * b + q * c - q * d) % R
    if Sp:
        h = (h * y + l0 * (1 - zp[0][0])) % R
        h = (h * y + l_last * (zp[-1][0] * zp[-1][0] - zp[-1][0])) % R
        for j in range(1, Sp):
            h = (h * y + l0 * (zp[j][0] - zp[j - 1][2])) % R
        for st in range(Sp):
            lhs, rhs = zp[st][1], zp[st][0]
            for col in range(2 * st, min(2 * st + 2, P)):
                v = adv[4 + L + col]
                lhs = lhs * (v + gamma + beta * sig[col]) % R
                rhs = rhs * (v + gamma + beta * pow(DELTA, col, R) * x) % R
            h = (h * y + l_active * (lhs - rhs)) % R
    for l in range(L):
        f, tb = (adv[4 + l] + beta) % R, (t + beta) % R
        z0, z1 = zl[l][0], zl[l][1]
        h = (h * y + l0 * z0) % R
        h = (h * y + l_last * z0) % R
        h = (h * y + l_active * (((z1 - z0) * tb + m[l]) * f - tb)) % R
    for i in range(S):
        inp, sh = (adv[4 + L + P + 2 * i] + beta) % R, (adv[4 + L + P + 2 * i + 1] + beta) % R
        z0, z1 = zs[i][0], zs[i][1]
        h = (h * y + l0 * (1 - z0)) % R
        h = (h * y + l_last * (z0 * z0 - z0)) % R
        h = (h * y + l_active * (sh * z1 - inp * z0)) % R
    return h, zh


@pytest.mark.parametrize("tamper", [False, True])
def test_quotient_identity_at_x(setup, zo, tamper):
    """the vanishing argument: h(x) (x^n - 1) = C(x) from the opened values
    (h(x) is the folded quotient's evaluation); a witness with one broken
    gate row makes h no polynomial, and the identity fails"""
    from zkprover.prover import BATCHED, TAMPER_GATE
    pk, pr = setup
    com, _, info = pr.prove(45, flags=BATCHED | (TAMPER_GATE if tamper else 0))
    ev_m, _ = pr.openings()
    ev = _ints(zo, ev_m)
    ch = [_ints(zo, np.array(info.challenges[c], dtype=np.uint64))[0] for c in range(4)]
    beta, gamma, y, x = ch
    g1 = G1(zo)
    sets, _ = _layout(pk.circuit, com, pk.fixed_commitments(), x, g1)
    evs = _evals_by_set(sets, ev)
    cx, zh = _constraints_at_x(pk.circuit, evs, x, beta, gamma, y, zo)
    hx = evs[0][len(sets[0]) - 1][0]
    assert (hx * zh % R == cx) == (not tamper)
