"""GPU parity: lookup multiplicity m(X) through the drop-in C-ABI entry points
prepare_table_lookup / calc_m (src/cuda/bn254_c.rs:281, :295; kernels
cuda/bn254.cu:1637-1912; caller src/lib.rs:404-504).

Small sizes are compared bit-for-bit with the oracle's restatement of the
reference's sort-and-scan (tables whose Montgomery word-0 low bits collide
heavily, duplicate values, several inputs accumulating onto a nonzero m).
At the prover's sizes (k = 16, 20, 22) the table is the range table of the
benchmark circuit (row j holds j mod 2^16) and the expected m is the closed
form: every input value v is counted on the last usable row of its residue
class.
"""
import ctypes

import numpy as np
import pytest

from test_oracle_lookup_m import _inputs_from, _raw_table, _with_duplicates

pytestmark = pytest.mark.gpu


def _dv():
    from zkprover import device as dv
    from zkprover._lib import call, lib
    return dv, call, lib


class _Lookup:
    """the buffers the reference's caller allocates (src/lib.rs:407-412)"""

    def __init__(self, dv, k):
        n = 1 << k
        self.k, self.n = k, n
        (self.tmp0, self.tmp1, self.index, self.sorted_index, self.start_offset,
         self.temp) = [dv.empty(n) for _ in range(6)]

    def prepare(self, dv, call, table_dev, u):
        call("prepare_table_lookup", dv.ptr(table_dev), dv.ptr(self.tmp0), dv.ptr(self.tmp1),
             dv.ptr(self.index), dv.ptr(self.sorted_index), dv.ptr(self.start_offset), dv.ptr(self.temp),
             32 * self.n, self.k, u, dv.stream_handle())

    def calc_m(self, dv, call, m_dev, table_dev, input_dev, u):
        call("calc_m", dv.ptr(m_dev), dv.ptr(table_dev), dv.ptr(input_dev), dv.ptr(self.sorted_index),
             dv.ptr(self.start_offset), dv.ptr(self.tmp0), dv.ptr(self.tmp1), dv.ptr(self.temp),
             32 * self.n, self.k, u, dv.stream_handle())
        return int(dv.from_dev(self.tmp1[:1])[0, 0] & np.uint64(0xFFFFFFFF)) != 0


def _gpu_m(m0, table, inputs, k, u):
    dv, call, _ = _dv()
    lk = _Lookup(dv, k)
    dm, dt = dv.to_dev(m0), dv.to_dev(table)
    lk.prepare(dv, call, dt, u)
    missing = False
    for inp in inputs:
        missing |= lk.calc_m(dv, call, dm, dt, dv.to_dev(inp), u)
    return dv.from_dev(dm), missing


@pytest.mark.parametrize("k,nkeys,ninputs,u_off", [(4, 3, 1, 2), (6, 5, 2, 6), (8, 40, 3, 6), (9, 1, 2, 6),
                                                   (10, 200, 2, 6), (12, 4096, 2, 6), (10, 7, 1, 0)])
def test_calc_m_matches_oracle(gpu, zo, k, nkeys, ninputs, u_off):
    n = 1 << k
    u = n - u_off
    table = _with_duplicates(_raw_table(n, k, nkeys, 100 + k), 200 + k)
    inputs = _inputs_from(table, u, ninputs, n, 300 + k) if u < n else [
        table[np.random.default_rng(k).integers(0, n, size=n)] for _ in range(ninputs)]
    m0 = zo.rand_fr(n, 400 + k)
    exp, emiss = zo.lookup_multiplicity(m0, table, inputs, k, u)
    got, miss = _gpu_m(m0, table, inputs, k, u)
    assert not emiss and not miss
    assert (got == exp).all()


def test_calc_m_all_equal_values(gpu, zo):
    """one value in every row: every insert folds into one slot"""
    k = 12
    n = 1 << k
    u = n - 6
    table = np.tile(zo.rand_fr(1, 5), (n, 1))
    inp = table.copy()
    exp, _ = zo.lookup_multiplicity(zo.fr_zeros(n), table, [inp], k, u)
    got, miss = _gpu_m(zo.fr_zeros(n), table, [inp], k, u)
    assert not miss and (got == exp).all()
    assert zo.from_mont(got[u - 1])[0] == u


def test_calc_m_missing_flag(gpu, zo):
    k = 8
    n = 1 << k
    u = n - 4
    table = _raw_table(n, k, 9, 8)
    only_past_u = table[n - 1].copy()
    inp = table[np.arange(n) % u].copy()
    inp[5] = only_past_u
    inp[9] = zo.rand_fr(1, 77)[0]  # in no row at all
    exp, emiss = zo.lookup_multiplicity(zo.fr_zeros(n), table, [inp], k, u)
    got, miss = _gpu_m(zo.fr_zeros(n), table, [inp], k, u)
    assert emiss and miss
    assert (got == exp).all()
    # a later call that finds everything clears the flag
    ok = table[np.arange(n) % u].copy()
    got2, miss2 = _gpu_m(zo.fr_zeros(n), table, [ok], k, u)
    assert not miss2


def test_prepare_replaces_previous_table(gpu, zo):
    """start_offset is reused across lookups (src/lib.rs:407-449): a second
    prepare must drop every entry of the first table"""
    dv, call, _ = _dv()
    k = 10
    n = 1 << k
    u = n - 6
    ta = _raw_table(n, k, 50, 1)
    tb = _raw_table(n, k, 50, 2)
    lk = _Lookup(dv, k)
    da, db = dv.to_dev(ta), dv.to_dev(tb)
    lk.prepare(dv, call, da, u)
    lk.prepare(dv, call, db, u)
    dm = dv.zeros(n)
    assert lk.calc_m(dv, call, dm, db, dv.to_dev(ta), u)  # table a's values are gone
    dm2 = dv.zeros(n)
    inp = tb[np.arange(n) % u]
    assert not lk.calc_m(dv, call, dm2, db, dv.to_dev(inp), u)
    exp, _ = zo.lookup_multiplicity(zo.fr_zeros(n), tb, [inp], k, u)
    assert (dv.from_dev(dm2) == exp).all()


def test_calc_m_degenerate_and_invalid(gpu, zo):
    dv, call, lib = _dv()
    k = 6
    n = 1 << k
    table = _raw_table(n, k, 9, 3)
    m0 = zo.rand_fr(n, 4)
    got, miss = _gpu_m(m0, table, [table], k, 0)  # no usable rows: m unchanged
    assert not miss and (got == m0).all()
    lk = _Lookup(dv, k)
    dt = dv.to_dev(table)
    s = dv.stream_handle()
    p = [dv.ptr(t) for t in (lk.tmp0, lk.tmp1, lk.index, lk.sorted_index, lk.start_offset, lk.temp)]
    hipErrorInvalidValue = 1
    assert lib().prepare_table_lookup(dv.ptr(dt), *p, 32 * n, 0, 0, s) == hipErrorInvalidValue
    assert lib().prepare_table_lookup(dv.ptr(dt), *p, 32 * n, k, n + 1, s) == hipErrorInvalidValue
    assert lib().calc_m(dv.ptr(dt), dv.ptr(dt), dv.ptr(dt), p[3], p[4], p[0], p[1], p[5], 32 * n, k, n + 1,
                        s) == hipErrorInvalidValue
    assert lib().calc_m(None, dv.ptr(dt), dv.ptr(dt), p[3], p[4], p[0], p[1], p[5], 32 * n, k, 4,
                        s) == hipErrorInvalidValue


@pytest.mark.parametrize("k", [16, 20, 22])
def test_calc_m_range_table_closed_form(gpu, zo, k):
    """benchmark range table (row j holds j mod 2^16) and two inputs of
    random range values: m = per-value counts on the last usable row of each
    residue class, zero elsewhere"""
    n = 1 << k
    u = n - 6
    tsize = 1 << 16
    vals = zo.to_mont(list(range(tsize)))
    table = vals[np.arange(n) % tsize]
    rng = np.random.default_rng(k)
    hi = min(tsize, u)  # values with a usable table row
    ivs = [rng.integers(0, hi, size=n) for _ in range(2)]
    got, miss = _gpu_m(zo.fr_zeros(n), table, [vals[iv] for iv in ivs], k, u)
    assert not miss
    v = np.arange(hi)
    last_row = v + tsize * ((u - 1 - v) // tsize)  # the last j < u with j = v mod tsize
    counts = np.zeros(hi, dtype=np.int64)
    for iv in ivs:
        counts += np.bincount(iv[:u], minlength=hi)
    hit = counts > 0
    assert (got[last_row[hit]] == zo.to_mont([int(c) for c in counts[hit]])).all()
    others = np.ones(n, dtype=bool)
    others[last_row[hit]] = False
    assert not got[others].any()
