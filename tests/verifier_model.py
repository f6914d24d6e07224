"""The halo2 verifier's side of a prover-core proof, restated in python
integers (test infrastructure).

Given the commitments a proof returns, the verifying key, the evaluations
and the challenges, it rebuilds what the verifier checks:
  * the polynomials of the opening with their evaluation order
    (src/lib.rs:1025-1078) and query order (src/lib.rs:1094-1176 with
    permutation_product_open / lookup_open / shuffle_open,
    src/multiopen.rs:564-660);
  * GWC: queries grouped by rotation, groups in Rotation order
    (construct_intermediate_sets, src/multiopen.rs:47-66), each batched as
    sum_i v^(len-1-i) C_i; with the setup's trapdoor tau,
    (tau - z) W_z = C_z - E_z G, and the verifier's pairing product check;
  * Shplonk: polynomials grouped by their rotation sets, sets ordered as
    BTreeSet<Rotation> (src/multiopen.rs:208-276), the pairing equation
    e([L], [tau]_2) = e(u [L] + F / z_0, G2);
  * the vanishing argument: h(x)(x^n - 1) equals the composed constraint
    (gates folded in y, permutation, logup, shuffle) from the evaluations.
"""
import numpy as np

R = 21888242871839275222246405745257275088548364400416034343698204186575808495617
# delta = 7^(2^28) mod r, the generator of the permutation column cosets
DELTA = 0x09226b6e22c6f0ca64ec26aad4c86e715b5f898e5e963f25870e56bbe533e9a2
LAST, CUR, NEXT = 0, 1, 2


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


def ints(zo, arr):
    return zo.from_mont(np.ascontiguousarray(arr, dtype=np.uint64).reshape(-1, 4))


def shape(circ):
    A, F, I, L, P, S = circ.advice, circ.fixed, circ.instance, circ.lookups, circ.perm_cols, circ.shuffles
    return dict(A=A, F=F, I=I, L=L, P=P, S=S, Sp=(P + 1) // 2, FG=F - 2, PA=P - I,
                col_lookup=4, col_perm=4 + L, col_shuffle=4 + L + P - I, col_fg=4 + L + P - I + 2 * S)


def split_commitments(circ, com):
    """the proof's commitments by role (transcript order)"""
    s = shape(circ)
    out, c = {}, 0
    for name, cnt in (("inst", s["I"]), ("adv", s["A"]), ("m", s["L"]), ("zp", s["Sp"]), ("zl", s["L"]),
                      ("zs", s["S"]), ("rnd", 1), ("pieces", 3)):
        out[name] = com[c:c + cnt]
        c += cnt
    out["opening"] = com[c:]
    return out


def layout(circ, com, vk, x, g1):
    """(commitments of the opened polynomials, evaluation order, query order);
    entries of the orders are (polynomial index, rotation)"""
    s = shape(circ)
    sc = split_commitments(circ, com)
    n = 1 << circ.k
    xn = pow(x, n, R)
    polys = []

    def add(c):
        polys.append(c)
        return len(polys) - 1

    h = add(g1.aff(g1.lin([(sc["pieces"][i], pow(xn, i, R)) for i in range(3)]))[0])
    inst = [add(c) for c in sc["inst"]]
    adv = [add(c) for c in sc["adv"]]
    fix = [add(vk[i]) for i in range(s["F"])]
    rnd = add(sc["rnd"][0])
    sig = [add(vk[s["F"] + i]) for i in range(s["P"])]
    zp = [add(c) for c in sc["zp"]]
    m = [add(c) for c in sc["m"]]
    zl = [add(c) for c in sc["zl"]]
    zs = [add(c) for c in sc["zs"]]
    ev = [(h, CUR)] + [(p, CUR) for p in inst + adv + fix + [rnd] + sig]
    for i, p in enumerate(zp):
        ev += [(p, CUR), (p, NEXT)] + ([(p, LAST)] if i != len(zp) - 1 else [])
    for l in range(s["L"]):
        ev += [(m[l], CUR), (zl[l], CUR), (zl[l], NEXT)]
    for p in zs:
        ev += [(p, CUR), (p, NEXT)]
    q = [(p, CUR) for p in inst + adv]
    for p in zp:
        q += [(p, CUR), (p, NEXT)]
    q += [(p, LAST) for p in list(reversed(zp))[1:]]
    for l in range(s["L"]):
        q += [(m[l], CUR), (zl[l], CUR), (zl[l], NEXT)]
    for p in zs:
        q += [(p, CUR), (p, NEXT)]
    q += [(p, CUR) for p in fix + sig] + [(h, CUR), (rnd, CUR)]
    roles = dict(h=h, inst=inst, adv=adv, fix=fix, rnd=rnd, sig=sig, zp=zp, m=m, zl=zl, zs=zs)
    return polys, ev, q, roles


def points(circ, x, zo):
    """the opening points of the rotations last, cur, next"""
    n = 1 << circ.k
    w = ints(zo, zo.root_of_unity(circ.k))[0]
    return [x * pow(w, n - (circ.blinding + 1), R) % R, x, x * w % R]


def eval_map(ev_order, evals):
    assert len(ev_order) == len(evals)
    return {k: e for k, e in zip(ev_order, evals)}


def gwc_groups(circ, com, vk, evals, ch, zo, g1):
    """per rotation group: (W affine row, z, C - E G Jacobian)"""
    x, v = ch["x"], ch["v"]
    polys, ev_order, queries, _ = layout(circ, com, vk, x, g1)
    em = eval_map(ev_order, evals)
    pts = points(circ, x, zo)
    W = split_commitments(circ, com)["opening"]
    out, gi = [], 0
    for r in (LAST, CUR, NEXT):
        grp = [p for p, rr in queries if rr == r]
        if not grp:
            continue
        k = len(grp)
        batch = g1.lin([(polys[grp[i]], pow(v, k - 1 - i, R)) for i in range(k)])
        e = sum(em[(grp[i], r)] * pow(v, k - 1 - i, R) for i in range(k)) % R
        out.append((W[gi], pts[r], g1.add(batch, g1.mul_aff(g1.g, -e))))
        gi += 1
    return out


def gwc_pairing_accepts(groups, tau, g1, u=0x5eed):
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


def gwc_accepts(circ, com, vk, evals, ch, tau, zo, g1, pairing=True):
    groups = gwc_groups(circ, com, vk, evals, ch, zo, g1)
    for w, z, rhs in groups:
        if not (g1.aff(g1.mul_aff(w, tau - z)) == g1.aff(rhs)).all():
            return False
    return gwc_pairing_accepts(groups, tau, g1) if pairing else True


def lagrange_at(xs, ys, t):
    acc = 0
    for k, (xk, yk) in enumerate(zip(xs, ys)):
        num, den = 1, 1
        for l, xl in enumerate(xs):
            if l != k:
                num = num * (t - xl) % R
                den = den * (xk - xl) % R
        acc = (acc + yk * num * pow(den, -1, R)) % R
    return acc


def shplonk_accepts(circ, com, vk, evals, ch, tau, zo, g1):
    """the Shplonk opening against the trapdoor in G1 and against the
    verifier's pairing equation; True only if both hold"""
    x, v, y, u = ch["x"], ch["v"], ch["y2"], ch["u"]
    polys, ev_order, queries, _ = layout(circ, com, vk, x, g1)
    em = eval_map(ev_order, evals)
    pts = points(circ, x, zo)
    rotval = [-(circ.blinding + 1), 0, 1]
    rots, order = {}, []
    for p, r in queries:
        if p not in rots:
            order.append(p)
            rots[p] = set()
        rots[p].add(r)
    keys = sorted({tuple(sorted(rotval[r] for r in rots[p])) for p in order})
    rot_of = {rotval[r]: r for r in range(3)}
    super_rots = sorted({r for p in order for r in rots[p]})
    hx, lx = split_commitments(circ, com)["opening"][:2]
    zt = 1
    for r in super_rots:
        zt = zt * (u - pts[r]) % R
    F, z0 = None, None
    for si, key in enumerate(keys):
        srots = [rot_of[rv] for rv in key]
        members = [p for p in order if tuple(sorted(rotval[r] for r in rots[p])) == key]
        k = len(members)
        vs = g1.lin([(polys[members[i]], pow(y, k - 1 - i, R)) for i in range(k)])
        ecomb = [sum(em[(members[i], r)] * pow(y, k - 1 - i, R) for i in range(k)) % R for r in srots]
        ru = lagrange_at([pts[r] for r in srots], ecomb, u)
        zs = 1
        for r in super_rots:
            if r not in srots:
                zs = zs * (u - pts[r]) % R
        if si == 0:
            z0 = zs
        term = g1.add(vs, g1.mul_aff(g1.g, -ru))
        term = g1.mul_aff(g1.aff(term)[0], zs * pow(v, len(keys) - 1 - si, R))
        F = term if F is None else g1.add(F, term)
    F = g1.add(F, g1.mul_aff(hx, -zt))
    rhs = g1.mul_aff(g1.aff(F)[0], pow(z0, -1, R))
    if not (g1.aff(g1.mul_aff(lx, tau - u)) == g1.aff(rhs)).all():
        return False
    import bn254_pairing as bp
    right = g1.aff(g1.add(g1.mul_aff(lx, u), rhs))
    return bp.pairing_check([(bp.g2_mul(bp.G2, tau), zo.g1a_to_ints(lx)),
                             (bp.g2_neg(bp.G2), zo.g1a_to_ints(right))])


def constraints_at_x(circ, com, vk, evals, ch, zo, g1):
    """(C(x), x^n - 1): the composed constraint as the prover folds it on the
    coset -- the gates y^FG q (a b + c - d) + ... + q (f_j a - e_j) (each a
    fold h = h y + gate), permutation p1, p2, one term per set, logup per
    lookup, shuffle per pair -- from the evaluations and the verifier's own
    l0, l_last, l_active at x"""
    s = shape(circ)
    x, beta, gamma, y = ch["x"], ch["beta"], ch["gamma"], ch["y"]
    _, ev_order, _, ro = layout(circ, com, vk, x, g1)
    em = eval_map(ev_order, evals)
    e = lambda p, r=CUR: em[(p, r)]
    n = 1 << circ.k
    u = n - (circ.blinding + 1)
    w = ints(zo, zo.root_of_unity(circ.k))[0]
    zh = (pow(x, n, R) - 1) % R
    ninv = pow(n, -1, R)

    def lag(i):
        wi = pow(w, i, R)
        return wi * zh % R * ninv % R * pow(x - wi, -1, R) % R

    l0, l_last = lag(0), lag(u)
    l_active = (1 - sum(lag(i) for i in range(u, n))) % R
    adv = [e(p) for p in ro["adv"]]
    inst = [e(p) for p in ro["inst"]]
    q, t = e(ro["fix"][0]), e(ro["fix"][1])
    f = [e(p) for p in ro["fix"][2:]]
    sig = [e(p) for p in ro["sig"]]
    a, b, c, d = adv[0:4]
    h = (q * a * b + q * c - q * d) % R
    for g in range(s["FG"]):
        h = (h * y + q * (f[g] * a - adv[s["col_fg"] + g])) % R
    Sp, P = s["Sp"], s["P"]
    vals = inst + [adv[s["col_perm"] + i] for i in range(s["PA"])]
    if Sp:
        zp = ro["zp"]
        h = (h * y + l0 * (1 - e(zp[0]))) % R
        h = (h * y + l_last * (e(zp[-1]) * e(zp[-1]) - e(zp[-1]))) % R
        for j in range(1, Sp):
            h = (h * y + l0 * (e(zp[j]) - e(zp[j - 1], LAST))) % R
        for st in range(Sp):
            lhs, rhs = e(zp[st], NEXT), e(zp[st])
            for col in range(2 * st, min(2 * st + 2, P)):
                lhs = lhs * (vals[col] + gamma + beta * sig[col]) % R
                rhs = rhs * (vals[col] + gamma + beta * pow(DELTA, col, R) * x) % R
            h = (h * y + l_active * (lhs - rhs)) % R
    for l in range(s["L"]):
        fl, tb = (adv[s["col_lookup"] + l] + beta) % R, (t + beta) % R
        z0, z1 = e(ro["zl"][l]), e(ro["zl"][l], NEXT)
        h = (h * y + l0 * z0) % R
        h = (h * y + l_last * z0) % R
        h = (h * y + l_active * (((z1 - z0) * tb + e(ro["m"][l])) * fl - tb)) % R
    for i in range(s["S"]):
        inp = (adv[s["col_shuffle"] + 2 * i] + beta) % R
        sh = (adv[s["col_shuffle"] + 2 * i + 1] + beta) % R
        z0, z1 = e(ro["zs"][i]), e(ro["zs"][i], NEXT)
        h = (h * y + l0 * (1 - z0)) % R
        h = (h * y + l_last * (z0 * z0 - z0)) % R
        h = (h * y + l_active * (sh * z1 - inp * z0)) % R
    return h, zh


def challenges(zo, info, open_ch=None):
    ch = {k: ints(zo, np.array(info.challenges[i], dtype=np.uint64))[0]
          for i, k in enumerate(("beta", "gamma", "y", "x"))}
    if open_ch is not None:
        v, y2, u = ints(zo, open_ch)
        ch.update(v=v, y2=y2, u=u)
    return ch


def quotient_identity_holds(circ, com, vk, evals, ch, zo, g1):
    cx, zh = constraints_at_x(circ, com, vk, evals, ch, zo, g1)
    return evals[0] * zh % R == cx
