"""CPU checks pinning the oracle's lookup multiplicity (prepare_table_lookup +
calc_m, cuda/bn254.cu:1637-1912) before the GPU is compared with it.

The oracle restates the reference's sort-and-scan; these tests check it
against what fixes the answer independently of that method:
* brute force over a dictionary (value -> largest usable row holding it);
* the logup identity sum_i 1/(f_i + beta) = sum_j m_j/(t_j + beta) over the
  usable rows, which any correct multiplicity satisfies;
* special cases the reference's kernels decide: the LARGEST usable row wins
  among equal table values (the scan runs downwards from the last sorted
  position, :1684-1692); rows at and past u neither match nor are counted;
  a usable input with no usable match is the reference's assert (:1694).
"""
import numpy as np
import pytest

R_TOP = 0x30644E72E131A029  # top 64-bit word of r


def _raw_table(n, k, nkeys, seed):
    """raw words below r whose word-0 low (k+2) bits collide heavily"""
    rng = np.random.default_rng(seed)
    t = rng.integers(0, 2**63, size=(n, 4), dtype=np.uint64)
    t[:, 3] &= np.uint64(0x0FFFFFFFFFFFFFFF)
    keys = rng.integers(0, nkeys, size=n, dtype=np.uint64)
    t[:, 0] = (t[:, 0] & ~np.uint64((1 << (k + 2)) - 1)) | keys
    return t


def _with_duplicates(t, seed, frac=0.3):
    rng = np.random.default_rng(seed)
    n = len(t)
    dst = rng.integers(0, n, size=int(n * frac))
    src = rng.integers(0, n, size=len(dst))
    t = t.copy()
    t[dst] = t[src]
    return t


def _brute(m0, table, inputs, u, zo):
    last = {}
    for j in range(u):
        last[table[j].tobytes()] = j  # ascending j: the largest row stays
    m = m0.copy()
    missing = False
    for inp in inputs:
        for i in range(u):
            j = last.get(inp[i].tobytes())
            if j is None:
                missing = True
                continue
            m[j] = zo.fr_add(m[j], zo.fr_from_u64(1))[0]
    return m, missing


def _inputs_from(table, u, count, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for c in range(count):
        pick = rng.integers(0, u, size=n)
        inp = table[pick].copy()
        # rows past u hold anything: they are neither matched nor counted
        inp[u:] = rng.integers(0, 2**62, size=(n - u, 4), dtype=np.uint64)
        out.append(inp)
    return out


@pytest.mark.parametrize("k,nkeys,ninputs", [(4, 3, 1), (6, 5, 2), (8, 40, 3), (9, 1, 2)])
def test_oracle_equals_brute_force(zo, k, nkeys, ninputs):
    n = 1 << k
    u = n - 6 if n > 16 else n - 2
    table = _with_duplicates(_raw_table(n, k, nkeys, k), k + 1)
    inputs = _inputs_from(table, u, ninputs, n, 7 * k)
    m0 = zo.rand_fr(n, 3 + k)
    got, miss = zo.lookup_multiplicity(m0, table, inputs, k, u)
    exp, emiss = _brute(m0, table, inputs, u, zo)
    assert not emiss and not miss
    assert (got == exp).all()


def test_oracle_logup_identity(zo):
    """sum over inputs of 1/(f+beta) equals sum over usable rows of m/(t+beta)"""
    k, tsize = 8, 40
    n = 1 << k
    u = n - 6
    vals = zo.to_mont(list(range(tsize)))
    table = vals[np.arange(n) % tsize]
    rng = np.random.default_rng(11)
    inputs = [vals[rng.integers(0, tsize, size=n)] for _ in range(2)]
    m, miss = zo.lookup_multiplicity(zo.fr_zeros(n), table, inputs, k, u)
    assert not miss
    beta = zo.rand_fr(1, 12)
    lhs = zo.fr_zeros(1)
    for inp in inputs:
        inv = zo.batch_invert(np.stack([zo.fr_add(inp[i], beta)[0] for i in range(u)]))
        for i in range(u):
            lhs = zo.fr_add(lhs, inv[i])
    tinv = zo.batch_invert(np.stack([zo.fr_add(table[j], beta)[0] for j in range(u)]))
    rhs = zo.fr_zeros(1)
    for j in range(u):
        rhs = zo.fr_add(rhs, zo.fr_mul(m[j], tinv[j]))
    assert (lhs == rhs).all()
    # every count lands on the last usable row of its residue class
    counts = np.zeros(n, dtype=np.int64)
    for inp in inputs:
        for i in range(u):
            v = zo.from_mont(inp[i])[0]
            counts[max(j for j in range(u) if j % tsize == v)] += 1
    assert zo.from_mont(m) == [int(c) for c in counts]


def test_oracle_largest_usable_row_wins(zo):
    k = 8
    n = 1 << k
    u = 250
    table = _raw_table(n, k, 200, 5)
    x = table[17].copy()
    for j in (3, 10, 200, 252):  # 252 is past u: never a match
        table[j] = x
    inp = np.tile(x, (n, 1))
    m, miss = zo.lookup_multiplicity(zo.fr_zeros(n), table, [inp], k, u)
    assert not miss
    nz = [j for j in range(n) if m[j].any()]
    assert nz == [200]
    assert zo.from_mont(m[200])[0] == u


def test_oracle_missing_values(zo):
    k = 6
    n = 1 << k
    u = n - 4
    table = _raw_table(n, k, 9, 8)
    only_past_u = table[n - 1].copy()
    table[n - 2] = only_past_u
    inp = table[np.arange(n) % u].copy()
    inp[5] = only_past_u  # present in the table, but only past u
    m, miss = zo.lookup_multiplicity(zo.fr_zeros(n), table, [inp], k, u)
    assert miss
    exp, emiss = _brute(zo.fr_zeros(n), table, [inp], u, zo)
    assert emiss and (m == exp).all()
    # rows past u of the input are never looked up
    inp2 = table[np.arange(n) % u].copy()
    inp2[u:] = only_past_u
    m2, miss2 = zo.lookup_multiplicity(zo.fr_zeros(n), table, [inp2], k, u)
    assert not miss2
    assert sum(zo.from_mont(m2)) == u
