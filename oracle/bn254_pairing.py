"""BN254 optimal-ate pairing over python integers -- TEST INFRASTRUCTURE ONLY.

The verifier side of the proof checks: a KZG opening is accepted when
e(W, [tau]_2 - z G2) = e(C - E G, G2).  Only tests/ use this module; the
product library never does.  Written from the textbook construction (no
code of the reference, which has no verifier): the tower is taken as one
degree-12 extension Fp[w] / (w^12 - 18 w^6 + 82), into which the sextic
twist maps G2 points (x, y) in Fp2 = Fp[u] / (u^2 + 1) as
(x' w^2, y' w^3) with x' = (x0 - 9 x1) + x1 w^6; the Miller loop runs over
6 t + 2 (t = 4965661367192848881), then adds pi(Q) and -pi^2(Q); the final
exponentiation is the plain power (p^12 - 1) / r.  Slow (about a second per
pairing) and simple on purpose.
"""
P = 21888242871839275222246405745257275088696311157297823662689037894645226208583
R = 21888242871839275222246405745257275088548364400416034343698204186575808495617
ATE_LOOP = 29793968203157093288  # 6 t + 2
B2 = None  # the twist's b, set below

# G2 generator (x = x0 + x1 u, y = y0 + y1 u)
G2_X = (10857046999023057135944570762232829481370756359578518086990519993285655852781,
        11559732032986387107991004021392285783925812861821192530917403151452391805634)
G2_Y = (8495653923123431417604973247489272438418190587263600148770280649306958101930,
        4082367875863433681332203403145435568316851327593401208105741076214120093531)


# ---- Fp2 ----------------------------------------------------------------------
def f2_add(a, b):
    return ((a[0] + b[0]) % P, (a[1] + b[1]) % P)


def f2_sub(a, b):
    return ((a[0] - b[0]) % P, (a[1] - b[1]) % P)


def f2_mul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def f2_inv(a):
    d = pow((a[0] * a[0] + a[1] * a[1]) % P, P - 2, P)
    return (a[0] * d % P, -a[1] * d % P)


def f2_scal(a, k):
    return (a[0] * k % P, a[1] * k % P)


B2 = f2_mul((3, 0), f2_inv((9, 1)))  # y^2 = x^3 + 3 / (9 + u)


def g2_on_curve(q):
    x, y = q
    return f2_sub(f2_mul(y, y), f2_add(f2_mul(f2_mul(x, x), x), B2)) == (0, 0)


def g2_add(p1, p2):
    if p1 is None:
        return p2
    if p2 is None:
        return p1
    (x1, y1), (x2, y2) = p1, p2
    if x1 == x2:
        if f2_add(y1, y2) == (0, 0):
            return None
        m = f2_mul(f2_scal(f2_mul(x1, x1), 3), f2_inv(f2_scal(y1, 2)))
    else:
        m = f2_mul(f2_sub(y2, y1), f2_inv(f2_sub(x2, x1)))
    x3 = f2_sub(f2_sub(f2_mul(m, m), x1), x2)
    return (x3, f2_sub(f2_mul(m, f2_sub(x1, x3)), y1))


def g2_mul(q, k):
    k %= R
    acc = None
    while k:
        if k & 1:
            acc = g2_add(acc, q)
        q = g2_add(q, q)
        k >>= 1
    return acc


def g2_neg(q):
    return None if q is None else (q[0], ((-q[1][0]) % P, (-q[1][1]) % P))


G2 = (G2_X, G2_Y)


# ---- Fp12 = Fp[w] / (w^12 - 18 w^6 + 82) ----------------------------------------
class F12:
    __slots__ = ("c",)

    def __init__(self, c):
        self.c = list(c) + [0] * (12 - len(c))

    @staticmethod
    def one():
        return F12([1])

    def __mul__(self, o):
        if isinstance(o, int):
            return F12([a * o % P for a in self.c])
        t = [0] * 23
        for i, a in enumerate(self.c):
            if a:
                for j, b in enumerate(o.c):
                    t[i + j] += a * b
        for d in range(22, 11, -1):  # w^12 = 18 w^6 - 82
            v = t[d]
            if v:
                t[d - 6] += 18 * v
                t[d - 12] -= 82 * v
        return F12([v % P for v in t[:12]])

    def __add__(self, o):
        return F12([(a + b) % P for a, b in zip(self.c, o.c)])

    def __sub__(self, o):
        return F12([(a - b) % P for a, b in zip(self.c, o.c)])

    def __neg__(self):
        return F12([(-a) % P for a in self.c])

    def __eq__(self, o):
        return self.c == o.c

    def inv(self):
        # extended Euclid over Fp[w] against the modulus
        lm, hm = [1] + [0] * 12, [0] * 13
        low, high = self.c + [0], [82, 0, 0, 0, 0, 0, P - 18, 0, 0, 0, 0, 0, 1]

        def deg(p):
            d = len(p) - 1
            while d and p[d] == 0:
                d -= 1
            return d

        while deg(low):
            r = _poly_div(high, low)
            r += [0] * (13 - len(r))
            nm, new = list(hm), list(high)
            for i in range(13):
                for j in range(13 - i):
                    nm[i + j] -= lm[i] * r[j]
                    new[i + j] -= low[i] * r[j]
            nm = [x % P for x in nm]
            new = [x % P for x in new]
            lm, low, hm, high = nm, new, lm, low
        inv0 = pow(low[0], P - 2, P)
        return F12([x * inv0 % P for x in lm[:12]])

    def __truediv__(self, o):
        return self * o.inv()

    def __pow__(self, e):
        acc, base = F12.one(), self
        while e:
            if e & 1:
                acc = acc * base
            base = base * base
            e >>= 1
        return acc


def _poly_div(a, b):
    def deg(p):
        d = len(p) - 1
        while d and p[d] == 0:
            d -= 1
        return d

    dega, degb = deg(a), deg(b)
    temp = list(a)
    o = [0] * len(a)
    inv_lead = pow(b[degb], P - 2, P)
    for i in range(dega - degb, -1, -1):
        o[i] = (o[i] + temp[degb + i] * inv_lead) % P
        for c in range(degb + 1):
            temp[c + i] = (temp[c + i] - o[i] * b[c]) % P
    return [x % P for x in o[:deg(o) + 1]]


W = F12([0, 1])


def _twist(q):
    (x0, x1), (y0, y1) = q
    nx = F12([(x0 - 9 * x1) % P] + [0] * 5 + [x1])
    ny = F12([(y0 - 9 * y1) % P] + [0] * 5 + [y1])
    return (nx * (W * W), ny * (W * W * W))


def _line(p1, p2, t):
    (x1, y1), (x2, y2), (xt, yt) = p1, p2, t
    if x1 != x2:
        m = (y2 - y1) / (x2 - x1)
        return m * (xt - x1) - (yt - y1)
    if y1 == y2:
        m = (x1 * x1 * 3) / (y1 * 2)
        return m * (xt - x1) - (yt - y1)
    return xt - x1


def _add12(p1, p2):
    if p1 is None:
        return p2
    (x1, y1), (x2, y2) = p1, p2
    if x1 == x2 and y1 == y2:
        m = (x1 * x1 * 3) / (y1 * 2)
    else:
        m = (y2 - y1) / (x2 - x1)
    x3 = m * m - x1 - x2
    return (x3, -(m * x3) + m * x1 - y1)


def _frob(p):
    return (p[0] ** P, p[1] ** P)


def miller_loop(q, p):
    """f_{6t+2,Q}(P) times the two Frobenius lines; q in G2 (Fp2 affine),
    p in G1 (affine ints); either None gives 1"""
    if q is None or p is None:
        return F12.one()
    Q = _twist(q)
    Pp = (F12([p[0]]), F12([p[1]]))
    Rr, f = Q, F12.one()
    for i in range(ATE_LOOP.bit_length() - 2, -1, -1):
        f = f * f * _line(Rr, Rr, Pp)
        Rr = _add12(Rr, Rr)
        if ATE_LOOP >> i & 1:
            f = f * _line(Rr, Q, Pp)
            Rr = _add12(Rr, Q)
    Q1 = _frob(Q)
    nQ2 = _frob(Q1)
    nQ2 = (nQ2[0], -nQ2[1])
    f = f * _line(Rr, Q1, Pp)
    Rr = _add12(Rr, Q1)
    f = f * _line(Rr, nQ2, Pp)
    return f


def final_exp(f):
    return f ** ((P ** 12 - 1) // R)


def pairing(q, p):
    return final_exp(miller_loop(q, p))


def pairing_check(pairs):
    """prod e(Q_i, P_i) == 1 (one final exponentiation for all)"""
    f = F12.one()
    for q, p in pairs:
        f = f * miller_loop(q, p)
    return final_exp(f) == F12.one()
