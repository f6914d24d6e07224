"""Writes the golden fixtures in this directory from the independent big-integer
model (tests/bigint_model.py), not from the C oracle or the HIP library.

The reference repository commits no golden vectors; its known-answer tests
are constructions (src/cuda/msm.rs:457-537, src/profile.rs:109-269), which
these fixtures instantiate at small sizes together with plain field, curve,
DFT and grand-product vectors.  Values are hex strings of canonical integers.

Run: python tests/golden/make_golden.py   (deterministic; seeded)
"""
import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import bigint_model as M  # noqa: E402


def h(v):
    return hex(v)


def pt(p):
    return None if p is None else [h(p[0]), h(p[1])]


def main():
    rnd = random.Random(20260101)
    out = {}
    # field arithmetic and Montgomery form, both fields
    for name, m in (("fr", M.R), ("fp", M.P)):
        pairs = [(0, 0), (1, m - 1), (m - 1, m - 1), (2, m // 2)] + [(rnd.randrange(m), rnd.randrange(m)) for _ in range(12)]
        out[f"{name}_ops"] = [{"a": h(a), "b": h(b), "add": h((a + b) % m), "sub": h((a - b) % m),
                               "mul": h(a * b % m), "inv_a": h(M.inv(a, m)), "mont_a": h(M.to_mont(a, m))}
                              for a, b in pairs]
    # curve: k*G, doubling, addition, negation
    ks = [1, 2, 3, M.R - 1, M.R - 2] + [rnd.randrange(M.R) for _ in range(6)]
    out["g1_mul"] = [{"k": h(k), "kG": pt(M.g1_mul(M.G, k))} for k in ks]
    p1, p2 = M.g1_mul(M.G, ks[6]), M.g1_mul(M.G, ks[7])
    out["g1_add"] = {"p": pt(p1), "q": pt(p2), "p_plus_q": pt(M.g1_add(p1, p2)), "two_p": pt(M.g1_add(p1, p1))}
    # small MSMs (random bases k_i G, including a repeated and an opposite point)
    for n in (1, 4, 9):
        kb = [rnd.randrange(1, M.R) for _ in range(n)]
        bases = [M.g1_mul(M.G, k) for k in kb]
        sc = [rnd.randrange(M.R) for _ in range(n)]
        if n == 9:
            bases[3] = bases[2]
            bases[5] = M.g1_neg(bases[4])
            sc[5] = sc[4]
            sc[8] = 0
        out[f"msm_{n}"] = {"bases": [pt(b) for b in bases], "scalars": [h(s) for s in sc],
                           "result": pt(M.msm(bases, sc))}
    # the reference's test_msm construction at n = 512 (src/cuda/msm.rs:477-511)
    cand = [rnd.randrange(M.R) for _ in range(256)]
    sc = [rnd.randrange(M.R) for _ in range(512)]
    k = sum(s * cand[i % 256] for i, s in enumerate(sc)) % M.R
    out["msm_kat_512"] = {"candidates": [h(c) for c in cand], "scalars": [h(s) for s in sc],
                          "result": pt(M.g1_mul(M.G, k))}
    # DFT / inverse DFT, n = 16, with the s_i = s_{i-1} + i input of src/profile.rs:211-218
    w = M.root_of_unity(4)
    s0 = rnd.randrange(M.R)
    xs = [(s0 + i * (i + 1) // 2) % M.R for i in range(16)]
    out["dft_16"] = {"omega": h(w), "input": [h(x) for x in xs], "output": [h(y) for y in M.dft(xs, w)]}
    # coset distribution (zeta^(i mod 3)) then DFT at 4n, n = 8
    zeta = pow(M.MULT_GEN, (M.R - 1) // 3, M.R)
    coeffs = [rnd.randrange(M.R) for _ in range(8)]
    ext = [c * pow(zeta, i % 3, M.R) % M.R for i, c in enumerate(coeffs)] + [0] * 24
    out["coset_8_to_32"] = {"zeta": h(zeta), "coeffs": [h(c) for c in coeffs], "omega": h(M.root_of_unity(5)),
                            "output": [h(y) for y in M.dft(ext, M.root_of_unity(5))]}
    # batch inversion with zeros
    xs = [rnd.randrange(M.R) for _ in range(12)] + [0, 0]
    out["batch_invert"] = {"input": [h(x) for x in xs], "output": [h(y) for y in M.batch_invert(xs)]}
    # logup z, two sets (src/lib.rs:640-703)
    n, u, beta = 16, 12, rnd.randrange(M.R)
    ins0 = [[rnd.randrange(M.R) for _ in range(n)] for _ in range(2)]
    table = [rnd.randrange(M.R) for _ in range(n)]
    m = [rnd.randrange(4) for _ in range(n)]
    s0 = [sum(M.inv(f[i] + beta, M.R) for f in ins0) % M.R for i in range(n)]
    z0, last0 = M.logup_z(s0, table, m, beta, 0, u)
    ins1 = [[rnd.randrange(M.R) for _ in range(n)] for _ in range(3)]
    s1 = [sum(M.inv(f[i] + beta, M.R) for f in ins1) % M.R for i in range(n)]
    z1 = M.exclusive_sum(s1, last0)
    out["logup_z"] = {"beta": h(beta), "u": u, "inputs0": [[h(v) for v in f] for f in ins0],
                      "inputs1": [[h(v) for v in f] for f in ins1], "table": [h(v) for v in table], "m": m,
                      "sum0": [h(v) for v in s0], "z0": [h(v) for v in z0], "last0": h(last0),
                      "z1": [h(v) for v in z1], "last1": h(z1[u])}
    # permutation grand product, one set of 2 columns, n = 16 (src/permutation.rs:44-128)
    log_n = 4
    n = 1 << log_n
    om = M.root_of_unity(log_n)
    delta = pow(M.MULT_GEN, 1 << M.TWO_ADICITY, M.R)
    beta, gamma = rnd.randrange(M.R), rnd.randrange(M.R)
    vals = [[rnd.randrange(M.R) for _ in range(n)] for _ in range(2)]
    sig = [[rnd.randrange(M.R) for _ in range(n)] for _ in range(2)]
    dstart = pow(delta, 2, M.R)
    ratios = [M.permutation_ratio_row([v[i] for v in vals], [s[i] for s in sig], beta, gamma, dstart, delta,
                                      pow(om, i, M.R)) for i in range(n)]
    out["permutation_z"] = {"beta": h(beta), "gamma": h(gamma), "delta": h(delta), "delta_start": h(dstart),
                            "omega": h(om), "values": [[h(v) for v in c] for c in vals],
                            "sigmas": [[h(v) for v in c] for c in sig], "z": [h(v) for v in M.exclusive_product(ratios)]}
    # shuffle grand product, group of 2, n = 16 (src/shuffle.rs:142-209)
    beta = rnd.randrange(M.R)
    inputs = [[rnd.randrange(M.R) for _ in range(n)] for _ in range(2)]
    shuffles = [[rnd.randrange(M.R) for _ in range(n)] for _ in range(2)]
    ratios = [M.shuffle_ratio_row([a[i] for a in inputs], [s[i] for s in shuffles], beta) for i in range(n)]
    out["shuffle_z"] = {"beta": h(beta), "inputs": [[h(v) for v in c] for c in inputs],
                        "shuffles": [[h(v) for v in c] for c in shuffles],
                        "z": [h(v) for v in M.exclusive_product(ratios)]}
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()
