"""ctypes wrapper over oracle/libzkoracle.so -- TEST INFRASTRUCTURE ONLY.

The CPU oracle is the checker the HIP product path is compared against.  Only
tests/, the smoke check in __graft_entry__.py and the CPU-baseline leg of
bench.py may import this module; the product library never calls it.

Arrays of field elements are numpy uint64 arrays of shape (n, 4): the exact
bytes of halo2's ``Fr([u64; 4])`` in Montgomery form.  G1 affine points are
(n, 8) uint64 (x limbs then y limbs) and Jacobian points (n, 12).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libzkoracle.so")
_lib = None

R_MODULUS = 21888242871839275222246405745257275088548364400416034343698204186575808495617
P_MODULUS = 21888242871839275222246405745257275088696311157297823662689037894645226208583


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def _ptrs(arrs):
    arr = (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    return arr


def fr_zeros(n):
    return np.zeros((n, 4), dtype=np.uint64)


def set_threads(n):
    lib().zo_set_threads(ctypes.c_int(n))


def get_threads():
    return lib().zo_get_threads()


# ---- conversions between python ints and limb arrays ----------------------
def int_to_limbs(v, nlimbs=4):
    return [(v >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(nlimbs)]


def limbs_to_int(row):
    return sum(int(x) << (64 * i) for i, x in enumerate(row))


def to_mont(ints):
    """python ints (canonical, < r) -> (n,4) Montgomery array"""
    a = np.array([int_to_limbs(v % R_MODULUS) for v in ints], dtype=np.uint64).reshape(-1, 4)
    out = fr_zeros(len(a))
    lib().zo_fr_to_mont_n(_p(out), _p(a), ctypes.c_size_t(len(a)))
    return out


def from_mont(arr):
    """(n,4) Montgomery array -> list of python ints"""
    arr = np.ascontiguousarray(arr, dtype=np.uint64).reshape(-1, 4)
    out = fr_zeros(len(arr))
    lib().zo_fr_from_mont_n(_p(out), _p(arr), ctypes.c_size_t(len(arr)))
    return [limbs_to_int(r) for r in out]


def to_mont_arr(canon):
    """(n,4) canonical limbs (< r) -> (n,4) Montgomery"""
    canon = np.ascontiguousarray(canon, dtype=np.uint64)
    out = fr_zeros(len(canon))
    lib().zo_fr_to_mont_n(_p(out), _p(canon), ctypes.c_size_t(len(canon)))
    return out


def arith_pattern(n, seed):
    """s_0 uniform from seed, s_i = s_{i-1} + i (src/profile.rs:211-218), Montgomery"""
    s0 = from_mont(rand_fr(1, seed))[0]
    i = np.arange(n, dtype=np.uint64)
    tri = (i * (i + np.uint64(1))) // np.uint64(2)  # < 2^64 for n <= 2^31
    base = np.array(int_to_limbs(s0), dtype=np.uint64)
    out = np.zeros((n, 4), dtype=np.uint64)
    lo = base[0] + tri
    carry = (lo < tri).astype(np.uint64)
    out[:, 0] = lo
    for k in range(1, 4):
        v = base[k] + carry
        carry = ((v < carry) & (carry > 0)).astype(np.uint64)
        out[:, k] = v
    # values < r + 2^64: subtract r once where needed (compare as python ints on the few rows)
    rl = int_to_limbs(R_MODULUS)
    ge = (out[:, 3] > rl[3]) | ((out[:, 3] == rl[3]) & ((out[:, 2] > rl[2]) | ((out[:, 2] == rl[2]) & (
        (out[:, 1] > rl[1]) | ((out[:, 1] == rl[1]) & (out[:, 0] >= rl[0]))))))
    for idx in np.nonzero(ge)[0]:
        out[idx] = int_to_limbs(limbs_to_int(out[idx]) - R_MODULUS)
    return to_mont_arr(out)


def fr_from_u64(v):
    out = fr_zeros(1)
    lib().zo_fr_from_u64(_p(out), ctypes.c_uint64(v))
    return out


def fr_mul(a, b):
    out = fr_zeros(1)
    lib().zo_fr_mul(_p(out), _p(np.ascontiguousarray(a)), _p(np.ascontiguousarray(b)))
    return out


def fr_add(a, b):
    out = fr_zeros(1)
    lib().zo_fr_add(_p(out), _p(np.ascontiguousarray(a)), _p(np.ascontiguousarray(b)))
    return out


def fr_sub(a, b):
    out = fr_zeros(1)
    lib().zo_fr_sub(_p(out), _p(np.ascontiguousarray(a)), _p(np.ascontiguousarray(b)))
    return out


def fr_inv(a):
    out = fr_zeros(1)
    lib().zo_fr_inv(_p(out), _p(np.ascontiguousarray(a)))
    return out


def fr_pow(a, e):
    out = fr_zeros(1)
    lib().zo_fr_pow(_p(out), _p(np.ascontiguousarray(a)), ctypes.c_uint64(e))
    return out


def fp_mul(a, b):
    out = fr_zeros(1)
    lib().zo_fp_mul(_p(out), _p(np.ascontiguousarray(a)), _p(np.ascontiguousarray(b)))
    return out


def fp_to_mont_int(v):
    a = np.array([int_to_limbs(v % P_MODULUS)], dtype=np.uint64)
    out = fr_zeros(1)
    lib().zo_fp_to_mont(_p(out), _p(a))
    return out


def fp_from_mont_int(arr):
    out = fr_zeros(1)
    lib().zo_fp_from_mont(_p(out), _p(np.ascontiguousarray(arr, dtype=np.uint64).reshape(1, 4)))
    return limbs_to_int(out[0])


# ---- seeded inputs ---------------------------------------------------------
def rand_fr(n, seed):
    out = fr_zeros(n)
    lib().zo_rand_fr(_p(out), ctypes.c_size_t(n), ctypes.c_uint64(seed))
    return out


def rand_fr_nonzero(n, seed):
    out = fr_zeros(n)
    lib().zo_rand_fr_nonzero(_p(out), ctypes.c_size_t(n), ctypes.c_uint64(seed))
    return out


def gen_bases_random(n, seed):
    out = np.zeros((n, 8), dtype=np.uint64)
    lib().zo_gen_bases_random(_p(out), ctypes.c_size_t(n), ctypes.c_uint64(seed))
    return out


def gen_bases_ap(n, seed):
    out = np.zeros((n, 8), dtype=np.uint64)
    lib().zo_gen_bases_ap(_p(out), ctypes.c_size_t(n), ctypes.c_uint64(seed))
    return out


# ---- G1 --------------------------------------------------------------------
def g1_generator():
    out = np.zeros((1, 8), dtype=np.uint64)
    lib().zo_g1_generator(_p(out))
    return out


def g1_mul(p_aff, s_mont):
    out = np.zeros((1, 12), dtype=np.uint64)
    lib().zo_g1_mul(_p(out), _p(np.ascontiguousarray(p_aff)), _p(np.ascontiguousarray(s_mont)))
    return out


def g1_add(p_jac, q_jac):
    out = np.zeros((1, 12), dtype=np.uint64)
    lib().zo_g1_add(_p(out), _p(np.ascontiguousarray(p_jac)), _p(np.ascontiguousarray(q_jac)))
    return out


def g1_add_mixed(p_jac, q_aff):
    out = np.zeros((1, 12), dtype=np.uint64)
    lib().zo_g1_add_mixed(_p(out), _p(np.ascontiguousarray(p_jac)), _p(np.ascontiguousarray(q_aff)))
    return out


def g1_dbl(p_jac):
    out = np.zeros((1, 12), dtype=np.uint64)
    lib().zo_g1_dbl(_p(out), _p(np.ascontiguousarray(p_jac)))
    return out


def g1_to_affine(p_jac):
    p_jac = np.ascontiguousarray(p_jac, dtype=np.uint64).reshape(-1, 12)
    out = np.zeros((len(p_jac), 8), dtype=np.uint64)
    lib().zo_g1_batch_to_affine(_p(out), _p(p_jac), ctypes.c_size_t(len(p_jac)))
    return out


def g1a_is_on_curve(p_aff):
    return bool(lib().zo_g1a_is_on_curve(_p(np.ascontiguousarray(p_aff))))


def g1a_to_ints(p_aff):
    """affine (1,8) Montgomery -> (x, y) canonical python ints"""
    p_aff = np.ascontiguousarray(p_aff, dtype=np.uint64).reshape(8)
    return (fp_from_mont_int(p_aff[:4].copy()), fp_from_mont_int(p_aff[4:].copy()))


# ---- MSM -------------------------------------------------------------------
def fr_inner_cycled(s, c):
    out = fr_zeros(1)
    lib().zo_fr_inner_cycled(_p(out), _p(np.ascontiguousarray(s)), ctypes.c_size_t(len(s)),
                             _p(np.ascontiguousarray(c)), ctypes.c_size_t(len(c)))
    return out


def msm_naive(points, scalars):
    out = np.zeros((1, 12), dtype=np.uint64)
    lib().zo_msm_naive(_p(out), _p(points), _p(scalars), ctypes.c_size_t(len(scalars)))
    return out


def msm(points, scalars, c=0):
    out = np.zeros((1, 12), dtype=np.uint64)
    lib().zo_msm(_p(out), _p(points), _p(scalars), ctypes.c_size_t(len(scalars)), ctypes.c_int(c))
    return out


# ---- NTT -------------------------------------------------------------------
def root_of_unity(log_n):
    out = fr_zeros(1)
    lib().zo_root_of_unity(_p(out), ctypes.c_int(log_n))
    return out


def ntt(a, log_n, omega, n_inv=None):
    a = np.array(a, dtype=np.uint64, copy=True, order="C")
    lib().zo_ntt(_p(a), ctypes.c_size_t(log_n), _p(np.ascontiguousarray(omega)),
                 _p(None if n_inv is None else np.ascontiguousarray(n_inv)))
    return a


def dft_naive(a, omega):
    out = fr_zeros(len(a))
    lib().zo_dft_naive(_p(out), _p(np.ascontiguousarray(a)), ctypes.c_size_t(len(a)),
                       _p(np.ascontiguousarray(omega)))
    return out


def expand_omega(n, omega):
    out = fr_zeros(n)
    lib().zo_expand_omega(_p(out), ctypes.c_size_t(n), _p(np.ascontiguousarray(omega)))
    return out


def extended_prepare(s, coset_powers, size, extended_size, to_coset):
    s = np.array(s, dtype=np.uint64, copy=True, order="C")
    assert len(s) >= extended_size
    lib().zo_extended_prepare(_p(s), _p(np.ascontiguousarray(coset_powers)),
                              ctypes.c_int(len(coset_powers) + 1), ctypes.c_size_t(size),
                              ctypes.c_size_t(extended_size), ctypes.c_int(to_coset))
    return s


# ---- batch inversion / logup ------------------------------------------------
def batch_invert(a):
    a = np.array(a, dtype=np.uint64, copy=True, order="C")
    lib().zo_batch_invert(_p(a), ctypes.c_size_t(len(a)))
    return a


def logup_sum_input_inv(sum_, inp, beta, init):
    sum_ = np.array(sum_, dtype=np.uint64, copy=True, order="C")
    lib().zo_logup_sum_input_inv(_p(sum_), _p(np.ascontiguousarray(inp)),
                                 _p(np.ascontiguousarray(beta)), ctypes.c_int(init),
                                 ctypes.c_size_t(len(inp)))
    return sum_


def eval_logup_z_compose(input_sum, table, multiplicity, beta, last_z, last_z_index):
    n = len(input_sum)
    z = fr_zeros(n)
    last_z = np.array(last_z, dtype=np.uint64, copy=True, order="C")
    lib().zo_eval_logup_z_compose(_p(z), _p(np.ascontiguousarray(input_sum)),
                                  _p(np.ascontiguousarray(table)),
                                  _p(np.ascontiguousarray(multiplicity)),
                                  _p(np.ascontiguousarray(beta)), _p(last_z),
                                  ctypes.c_int(last_z_index), ctypes.c_size_t(n))
    return z, last_z


def eval_logup_z_pure(z_in, last_z, last_z_index):
    z = np.array(z_in, dtype=np.uint64, copy=True, order="C")
    last_z = np.array(last_z, dtype=np.uint64, copy=True, order="C")
    lib().zo_eval_logup_z_pure(_p(z), _p(last_z), ctypes.c_int(last_z_index),
                               ctypes.c_size_t(len(z)))
    return z, last_z


# ---- lookup multiplicity m(X) -------------------------------------------------
def lookup_multiplicity(m, table, inputs, k, unusable_rows_start):
    """Restates prepare_table_lookup + calc_m (cuda/bn254.cu:1637-1912) for
    each input of `inputs` in turn; returns (new m, missing).

    prepare (bn254.cu:1867-1912): key[j] = (j >= u) << (k+2) | (low k+2 bits
    of Montgomery word 0 of table[j]) (:1645); a stable sort of (key, j)
    (cub SortPairs, :1880-1897); start_offset[key] = the last sorted position
    holding key (:1649-1669).  calc_m (:1794-1865): for a usable input row,
    scan sorted positions downwards from start_offset[its key] and take the
    first table row with equal words (:1671-1700) -- so the largest usable
    row holding that value; rows past u never match since their keys sort
    above every usable key.  The matched rows are then counted and
    m[j] += count (:1702-1792).  `missing` is True when some usable input has
    no match (the reference's device assert, :1694).  Python loops: small n.
    """
    n = 1 << k
    table = np.ascontiguousarray(table, dtype=np.uint64)
    m = np.array(m, dtype=np.uint64, copy=True, order="C")
    mask = (1 << (k + 2)) - 1
    rows = np.arange(n, dtype=np.uint64)
    low = (table[:, 0] & np.uint64(0xFFFFFFFF)).astype(np.uint64)
    key = ((rows >= np.uint64(unusable_rows_start)).astype(np.uint64) << np.uint64(k + 2)) | (low & np.uint64(mask))
    order = np.argsort(key, kind="stable")
    skey = key[order]
    start_offset = {}
    for pos in range(n):
        start_offset[int(skey[pos])] = pos  # the last position wins
    missing = False
    for inp in inputs:
        inp = np.ascontiguousarray(inp, dtype=np.uint64)
        counts = {}
        for i in range(unusable_rows_start):
            pos = start_offset.get(int(inp[i, 0]) & mask, -1)
            while pos >= 0:
                j = int(order[pos])
                if (inp[i] == table[j]).all():
                    counts[j] = counts.get(j, 0) + 1
                    break
                pos -= 1
            else:
                missing = True
        for j, c in counts.items():
            if j < unusable_rows_start:
                m[j] = fr_add(m[j], fr_from_u64(c))[0]
    return m, missing


# ---- grand products ----------------------------------------------------------
def permutation_z_set(values, sigmas, beta, gamma, delta_start, delta, omega):
    n = len(values[0])
    z = fr_zeros(n)
    lib().zo_permutation_z_set(_p(z), _ptrs(values), _ptrs(sigmas), ctypes.c_int(len(values)),
                               _p(beta), _p(gamma), _p(delta_start), _p(delta), _p(omega),
                               ctypes.c_size_t(n))
    return z


def permutation_chain(zs, u):
    zs = [np.array(z, dtype=np.uint64, copy=True, order="C") for z in zs]
    lib().zo_permutation_chain(_ptrs(zs), ctypes.c_int(len(zs)), ctypes.c_size_t(u))
    return zs


def shuffle_z(inputs, shuffles, beta):
    n = len(inputs[0])
    z = fr_zeros(n)
    lib().zo_shuffle_z(_p(z), _ptrs(inputs), _ptrs(shuffles), ctypes.c_int(len(inputs)),
                       _p(beta), ctypes.c_size_t(n))
    return z


# ---- eval_h per-row formulas ---------------------------------------------------
def field_op(l, l_rot, l_c, r, r_rot, r_c, n, op):
    res = fr_zeros(n)
    lib().zo_field_op(_p(res), _p(l), ctypes.c_int(l_rot), _p(l_c), _p(r), ctypes.c_int(r_rot),
                      _p(r_c), ctypes.c_size_t(n), ctypes.c_int(op))
    return res


def field_op_batch_mul_sum(res, terms, n):
    """terms: list of (coeff (1,4), [(array, rot), ...])"""
    res = np.array(res, dtype=np.uint64, copy=True, order="C")
    vals, rots, coeffs, offs = [], [], [], [0]
    for coeff, cols in terms:
        coeffs.append(np.ascontiguousarray(coeff).reshape(4))
        for arr, rot in cols:
            vals.append(arr)
            rots.append(rot)
        offs.append(len(vals))
    coeffs = np.array(coeffs, dtype=np.uint64).reshape(-1, 4)
    rots_a = (ctypes.c_int * max(1, len(rots)))(*rots)
    offs_a = (ctypes.c_int * len(offs))(*offs)
    lib().zo_field_op_batch_mul_sum(_p(res), _ptrs(vals) if vals else None, rots_a, _p(coeffs),
                                    offs_a, ctypes.c_int(len(terms)), ctypes.c_size_t(n))
    return res


def field_mul_zip(buf, coeff):
    buf = np.array(buf, dtype=np.uint64, copy=True, order="C")
    lib().zo_field_mul_zip(_p(buf), _p(np.ascontiguousarray(coeff)), ctypes.c_size_t(len(coeff)),
                           ctypes.c_size_t(len(buf)))
    return buf


def _copy(a):
    return np.array(a, dtype=np.uint64, copy=True, order="C")


def permutation_eval_h_p1(res, first, last, l0, l_last, y):
    res = _copy(res)
    lib().zo_permutation_eval_h_p1(_p(res), _p(first), _p(last), _p(l0), _p(l_last), _p(y),
                                   ctypes.c_size_t(len(res)))
    return res


def permutation_eval_h_p2(res, sets, l0, y, rot):
    res = _copy(res)
    lib().zo_permutation_eval_h_p2(_p(res), _ptrs(sets), _p(l0), _p(y), ctypes.c_int(len(sets)),
                                   ctypes.c_int(rot), ctypes.c_size_t(len(res)))
    return res


def permutation_eval_h_l(res, beta, gamma, p):
    res = _copy(res)
    lib().zo_permutation_eval_h_l(_p(res), _p(beta), _p(gamma), _p(p), ctypes.c_size_t(len(res)))
    return res


def logup_eval_h(res, input_product, input_product_sum, table, multiplicity, z_first, z_last,
                 l0, l_last, l_active_row, y, rot):
    res = _copy(res)
    lib().zo_logup_eval_h(_p(res), _p(input_product), _p(input_product_sum), _p(table),
                          _p(multiplicity), _p(z_first), _p(z_last), _p(l0), _p(l_last),
                          _p(l_active_row), _p(y), ctypes.c_int(rot), ctypes.c_size_t(len(res)))
    return res


def logup_eval_h_extra_inputs(res, input_product, input_product_sum, z, l_active_row, y, rot):
    res = _copy(res)
    lib().zo_logup_eval_h_extra_inputs(_p(res), _p(input_product), _p(input_product_sum), _p(z),
                                       _p(l_active_row), _p(y), ctypes.c_int(rot),
                                       ctypes.c_size_t(len(res)))
    return res


def field_sum(vs, cs, rots, omegas):
    """cuda/bn254.cu:496-540; cs entries (1, 4) or None"""
    n = len(vs[0])
    res = fr_zeros(n)
    keep = [None if c is None else np.ascontiguousarray(c, dtype=np.uint64) for c in cs]
    cptr = (ctypes.c_void_p * len(cs))(*[None if c is None else c.ctypes.data for c in keep])
    rot = (ctypes.c_int * len(rots))(*rots)
    lib().zo_field_sum(_p(res), _ptrs(vs), cptr, rot, _p(np.ascontiguousarray(omegas)), ctypes.c_int(len(vs)),
                       ctypes.c_size_t(n))
    return res


def lookup_eval_h(res, inp, table, pin, ptab, z, l0, l_last, l_active_row, y, beta, gamma, rot):
    res = _copy(res)
    lib().zo_lookup_eval_h(_p(res), _p(inp), _p(table), _p(pin), _p(ptab), _p(z), _p(l0), _p(l_last),
                           _p(l_active_row), _p(y), _p(beta), _p(gamma), ctypes.c_int(rot),
                           ctypes.c_size_t(len(res)))
    return res


def logup_eval_h_inputs_product_sum(sets):
    n = len(sets[0])
    prod, psum = fr_zeros(n), fr_zeros(n)
    lib().zo_logup_eval_h_inputs_product_sum(_p(prod), _p(psum), _ptrs(sets), ctypes.c_int(len(sets)),
                                             ctypes.c_size_t(n))
    return prod, psum


def shuffle_eval_h(res, inp, table, z, l0, l_last, l_active_row, y, rot):
    res = _copy(res)
    lib().zo_shuffle_eval_h(_p(res), _p(inp), _p(table), _p(z), _p(l0), _p(l_last),
                            _p(l_active_row), _p(y), ctypes.c_int(rot), ctypes.c_size_t(len(res)))
    return res


def shuffle_eval_h_v2(res, inputs, tables, betas, z, l0, l_last, l_active_row, y, rot):
    res = _copy(res)
    lib().zo_shuffle_eval_h_v2(_p(res), _ptrs(inputs), _ptrs(tables), _p(betas),
                               ctypes.c_int(len(inputs)), _p(z), _p(l0), _p(l_last),
                               _p(l_active_row), _p(y), ctypes.c_int(rot), ctypes.c_size_t(len(res)))
    return res


# ---- vanishing construction (composition of the primitives above) ----------
def vanishing_construct(h, t_evals, log_n, log_ext, ext_omega_inv, ext_divisor, coset_powers, g, q, xn):
    """src/eval_h.rs:464-583: h *= t_evals[i mod |t|]; extended iNTT with the
    divisor; coset un-distribution (powers [g_coset_inv, g_coset], whole
    extended domain); commitments of the q pieces of n against g; h_pieces by
    Horner in xn.  Returns (coeffs (N), commitments (q affine), h_pieces (n))."""
    n = 1 << log_n
    N = 1 << log_ext
    x = field_mul_zip(h, t_evals)
    x = ntt(x, log_ext, ext_omega_inv, ext_divisor)
    x = extended_prepare(x, coset_powers, n, N, 1)
    commits = np.concatenate([g1_to_affine(msm(g, np.ascontiguousarray(x[i * n:(i + 1) * n])))
                              for i in range(q)])
    last = np.array(x[(q - 1) * n:q * n], copy=True)
    for i in range(q - 2, -1, -1):
        last = field_op(last, 0, xn, np.ascontiguousarray(x[i * n:(i + 1) * n]), 0, None, n, 0)
    return x, commits, last


# ---- multi-open -----------------------------------------------------------------
def x_table(x, k):
    """[x, x^2, x^4, ..., x^(2^(k-1))] as src/eval_poly.rs:44-53 builds it"""
    out = [np.ascontiguousarray(x, dtype=np.uint64).reshape(1, 4)]
    for _ in range(1, max(k, 1)):
        out.append(fr_mul(out[-1], out[-1]))
    return np.concatenate(out)


def poly_eval(p, xt):
    """cuda/bn254.cu:1608-1633 (rounds of <= 4 bits); returns p(x) as (1, 4)"""
    p = np.ascontiguousarray(p, dtype=np.uint64)
    res = fr_zeros(len(p))
    lib().zo_poly_eval(_p(res), _p(p), _p(np.ascontiguousarray(xt, dtype=np.uint64)), ctypes.c_size_t(len(p)))
    return res[:1].copy()


def kate_divide(poly, z):
    """src/multiopen.rs:145-159: in-place division by (X - z)"""
    out = np.array(poly, dtype=np.uint64, copy=True, order="C")
    lib().zo_kate_divide(_p(out), _p(np.ascontiguousarray(z, dtype=np.uint64)), ctypes.c_size_t(len(out)))
    return out


def shplonk_h_x_merge(res, v, values, omegas, diff):
    out = np.array(res, dtype=np.uint64, copy=True, order="C")
    diff = np.ascontiguousarray(diff, dtype=np.uint64).reshape(-1, 4)
    lib().zo_shplonk_h_x_merge(_p(out), _p(np.ascontiguousarray(v)), _p(np.ascontiguousarray(values)),
                               _p(np.ascontiguousarray(omegas)), _p(diff) if len(diff) else None,
                               ctypes.c_int(len(diff)), ctypes.c_size_t(len(out)))
    return out


def shplonk_h_x_div_points(values, omegas, points):
    out = np.array(values, dtype=np.uint64, copy=True, order="C")
    points = np.ascontiguousarray(points, dtype=np.uint64).reshape(-1, 4)
    lib().zo_shplonk_h_x_div_points(_p(out), _p(np.ascontiguousarray(omegas)), _p(points),
                                    ctypes.c_int(len(points)), ctypes.c_size_t(len(out)))
    return out
