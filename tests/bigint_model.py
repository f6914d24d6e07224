"""Independent pure-Python model of the BN254 arithmetic (test infrastructure).

Written from the mathematics (field definitions, short-Weierstrass affine
group law, the DFT definition, running sums / products), not from the C
oracle, so that agreement between the two pins the oracle.  Values are
canonical Python ints; Montgomery form only appears in the byte helpers.
"""
R = 21888242871839275222246405745257275088548364400416034343698204186575808495617
P = 21888242871839275222246405745257275088696311157297823662689037894645226208583
MONT = 1 << 256
G = (1, 2)        # BN254 G1 generator
B_COEFF = 3       # y^2 = x^3 + 3
TWO_ADICITY = 28
MULT_GEN = 7      # multiplicative generator of Fr*


# ---- fields -------------------------------------------------------------------
def inv(a, m):
    return pow(a, -1, m) if a % m else 0


def to_mont(v, m=R):
    return v * MONT % m


def from_mont(v, m=R):
    return v * pow(MONT, -1, m) % m


def limbs(v):
    return [(v >> (64 * i)) & (2 ** 64 - 1) for i in range(4)]


def unlimbs(ls):
    return sum(int(x) << (64 * i) for i, x in enumerate(ls))


def root_of_unity(log_n):
    w = pow(MULT_GEN, (R - 1) >> TWO_ADICITY, R)
    return pow(w, 1 << (TWO_ADICITY - log_n), R)


# ---- G1 (affine, None = identity) ------------------------------------------------
def on_curve(pt):
    if pt is None:
        return True
    x, y = pt
    return (y * y - x * x * x - B_COEFF) % P == 0


def g1_add(a, b):
    if a is None:
        return b
    if b is None:
        return a
    (x1, y1), (x2, y2) = a, b
    if x1 == x2:
        if (y1 + y2) % P == 0:
            return None
        lam = 3 * x1 * x1 * inv(2 * y1, P) % P
    else:
        lam = (y2 - y1) * inv(x2 - x1, P) % P
    x3 = (lam * lam - x1 - x2) % P
    return (x3, (lam * (x1 - x3) - y1) % P)


def g1_neg(a):
    return None if a is None else (a[0], (-a[1]) % P)


def g1_mul(pt, k):
    acc, base = None, pt
    k %= R
    while k:
        if k & 1:
            acc = g1_add(acc, base)
        base = g1_add(base, base)
        k >>= 1
    return acc


def msm(points, scalars):
    acc = None
    for pt, s in zip(points, scalars):
        acc = g1_add(acc, g1_mul(pt, s))
    return acc


# ---- polynomial / vector primitives ------------------------------------------------
def dft(xs, w):
    n = len(xs)
    return [sum(x * pow(w, i * j, R) for i, x in enumerate(xs)) % R for j in range(n)]


def batch_invert(xs):
    return [inv(x, R) for x in xs]


def exclusive_sum(xs, start=0):
    out, acc = [], start
    for x in xs:
        out.append(acc)
        acc = (acc + x) % R
    return out


def exclusive_product(xs, start=1):
    out, acc = [], start
    for x in xs:
        out.append(acc)
        acc = acc * x % R
    return out


# ---- the reference's argument formulas, stated on canonical ints ---------------------
def logup_z(input_sets_sum, table, m, beta, carry, u):
    """z[i] = carry + sum_{j<i} (S_j - m_j / (t_j + beta)); returns (z, z[u])"""
    terms = [(s - mm * inv(t + beta, R)) % R for s, t, mm in zip(input_sets_sum, table, m)]
    z = exclusive_sum(terms, carry)
    return z, z[u]


def permutation_ratio_row(vals, sigmas, beta, gamma, delta_start, delta, omega_i):
    num = den = 1
    d = delta_start
    for v, s in zip(vals, sigmas):
        num = num * (v + beta * d * omega_i + gamma) % R
        den = den * (v + beta * s + gamma) % R
        d = d * delta % R
    return num * inv(den, R) % R


def shuffle_ratio_row(inputs, shuffles, beta):
    num = den = 1
    for j, (a, s) in enumerate(zip(inputs, shuffles)):
        bj = pow(beta, j + 1, R)
        num = num * (a + bj) % R
        den = den * (s + bj) % R
    return num * inv(den, R) % R
