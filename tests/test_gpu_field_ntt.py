"""GPU parity: BN254 field ops, NTT / iNTT, twiddle tables, coset preparation.

Every GPU output is compared bit-for-bit with the CPU oracle on the same seeded
inputs.  NTT cases follow the reference's own test (src/profile.rs:174-269):
s_0 random, s_i = s_{i-1} + i, forward NTT equal to best_fft (the DFT) and the
iNTT with divisor 2^-k round-tripping exactly.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

R = 21888242871839275222246405745257275088548364400416034343698204186575808495617
P = 21888242871839275222246405745257275088696311157297823662689037894645226208583


def _edge_fr(zo):
    vals = [0, 1, 2, R - 1, R - 2, (1 << 253), R // 2, R // 3]
    return zo.to_mont(vals)


def test_fr_ops_match_oracle(gpu, zo):
    from zkprover import device as dv
    from zkprover._lib import call
    a = np.concatenate([_edge_fr(zo), zo.rand_fr(4096, 11)])
    b = np.concatenate([_edge_fr(zo)[::-1].copy(), zo.rand_fr(4096, 12)])
    n = len(a)
    da, db, do = dv.to_dev(a), dv.to_dev(b), dv.empty(n)
    ai, bi = zo.from_mont(a), zo.from_mont(b)
    expect = {
        0: [(x + y) % R for x, y in zip(ai, bi)],
        1: [(x * y) % R for x, y in zip(ai, bi)],
        2: [(x - y) % R for x, y in zip(ai, bi)],
        3: [pow(x, -1, R) if x else 0 for x in ai],
        4: [(x * x) % R for x in ai],
        5: [(-x) % R for x in ai],
    }
    for op, exp in expect.items():
        call("zk_fr_op", dv.ptr(do), dv.ptr(da), dv.ptr(db), op, n, dv.stream_handle())
        got = zo.from_mont(dv.from_dev(do))
        assert got == exp, f"op {op}"


def test_fp_mul_inv(gpu, zo):
    from zkprover import device as dv
    from zkprover._lib import call
    rng = np.random.default_rng(5)
    xs = [int.from_bytes(rng.bytes(32), "little") % P for _ in range(1024)] + [0, 1, P - 1]
    ys = xs[::-1]
    a = np.concatenate([zo.fp_to_mont_int(x) for x in xs])
    b = np.concatenate([zo.fp_to_mont_int(y) for y in ys])
    n = len(xs)
    da, db, do = dv.to_dev(a), dv.to_dev(b), dv.empty(n)
    call("zk_fp_op", dv.ptr(do), dv.ptr(da), dv.ptr(db), 1, n, dv.stream_handle())
    got = dv.from_dev(do)
    assert [zo.fp_from_mont_int(got[i]) for i in range(n)] == [x * y % P for x, y in zip(xs, ys)]
    call("zk_fp_op", dv.ptr(do), dv.ptr(da), None, 3, n, dv.stream_handle())
    got = dv.from_dev(do)
    assert [zo.fp_from_mont_int(got[i]) for i in range(n)] == [pow(x, -1, P) if x else 0 for x in xs]


def _omegas_dev(zo, log_n, omega):
    """mirror of generate_omega_buffers (src/cuda/ntt.rs:15-34)"""
    from zkprover import device as dv
    from zkprover._lib import call
    n = 1 << log_n
    init = np.zeros((n, 4), dtype=np.uint64)
    init[0] = zo.fr_from_u64(1)[0]
    if n > 1:
        init[1] = omega[0]
    buf = dv.to_dev(init)
    call("expand_omega_buffer", dv.ptr(buf), log_n, 1)
    return buf


def _gpu_ntt(zo, x, log_n, omega, n_inv=None):
    import ctypes
    from zkprover import device as dv
    from zkprover._lib import call
    om = _omegas_dev(zo, log_n, omega)
    buf, tmp = dv.to_dev(x), dv.empty(len(x))
    ninv = dv.to_dev(n_inv) if n_inv is not None else None
    swap = ctypes.c_bool(False)
    call("ntt", dv.ptr(buf), dv.ptr(tmp), None, dv.ptr(om), dv.ptr(ninv), log_n, 8,
         ctypes.addressof(swap), dv.stream_handle())
    return dv.from_dev(tmp if swap.value else buf)


def test_expand_omega_matches_powers(gpu, zo):
    from zkprover import device as dv
    for log_n in (1, 5, 12, 17):
        w = zo.root_of_unity(log_n)
        got = dv.from_dev(_omegas_dev(zo, log_n, w))
        assert (got == zo.expand_omega(1 << log_n, w)).all(), log_n


@pytest.mark.parametrize("log_n", list(range(1, 13)) + [14, 16, 18])
def test_ntt_matches_oracle(gpu, zo, log_n):
    n = 1 << log_n
    x = zo.rand_fr(n, 100 + log_n)
    w = zo.root_of_unity(log_n)
    got = _gpu_ntt(zo, x, log_n, w)
    exp = zo.ntt(x, log_n, w)
    assert (got == exp).all()
    if log_n <= 8:
        assert (exp == zo.dft_naive(x, w)).all()


@pytest.mark.parametrize("log_n", [20, 22])
def test_ntt_reference_pattern_roundtrip(gpu, zo, log_n):
    # s_0 random, s_i = s_{i-1} + i  (src/profile.rs:211-218)
    n = 1 << log_n
    x = zo.arith_pattern(n, 7)
    w = zo.root_of_unity(log_n)
    got = _gpu_ntt(zo, x, log_n, w)
    exp = zo.ntt(x, log_n, w)
    assert (got == exp).all()
    w_inv = zo.fr_inv(w)
    n_inv = zo.fr_inv(zo.fr_from_u64(n))
    back = _gpu_ntt(zo, got, log_n, w_inv, n_inv)
    assert (back == x).all()


def test_intt_roundtrip_2_24(gpu, zo):
    log_n = 24
    n = 1 << log_n
    x = zo.rand_fr(n, 99)
    w = zo.root_of_unity(log_n)
    y = _gpu_ntt(zo, x, log_n, w)
    exp = zo.ntt(x, log_n, w)
    assert (y == exp).all()
    back = _gpu_ntt(zo, y, log_n, zo.fr_inv(w), zo.fr_inv(zo.fr_from_u64(n)))
    assert (back == x).all()


@pytest.mark.parametrize("log_n,ext", [(10, 2), (16, 2), (20, 2), (22, 2)])
def test_coset_extend_matches_oracle(gpu, zo, log_n, ext):
    """extended_prepare (zero-pad + zeta^(i mod 3) scaling) then NTT at 2^(k+ext)
    (src/eval_h.rs:250-337, cuda/bn254.cu:641-662, :1232-1253).  Since
    zeta^3 = 1, zeta^(i mod 3) = zeta^i and output j is p(zeta w_N^j): a few
    outputs are also checked as evaluations of the coefficients there."""
    import ctypes
    from zkprover import device as dv
    from zkprover._lib import call
    n, N = 1 << log_n, 1 << (log_n + ext)
    coeffs = zo.rand_fr(n, 300 + log_n)
    zeta = zo.to_mont([pow(7, (R - 1) // 3, R)])
    zeta2 = zo.fr_mul(zeta, zeta)
    cp = np.concatenate([zeta, zeta2])
    host = np.zeros((N, 4), dtype=np.uint64)
    host[:n] = coeffs
    s = dv.to_dev(host)
    dcp = dv.to_dev(cp)
    call("extended_prepare", dv.ptr(s), dv.ptr(dcp), 3, n, N, 0, dv.stream_handle())
    exp_prep = zo.extended_prepare(host, cp, n, N, 0)
    assert (dv.from_dev(s) == exp_prep).all()
    w = zo.root_of_unity(log_n + ext)
    om = _omegas_dev(zo, log_n + ext, w)
    tmp = dv.empty(N)
    swap = ctypes.c_bool(False)
    call("ntt", dv.ptr(s), dv.ptr(tmp), None, dv.ptr(om), None, log_n + ext, 8,
         ctypes.addressof(swap), dv.stream_handle())
    got = dv.from_dev(tmp if swap.value else s)
    assert (got == zo.ntt(exp_prep, log_n + ext, w)).all()
    wi = zo.from_mont(w)[0]
    for j in (0, 1, N // 3, N - 1):
        x = zo.to_mont([pow(7, (R - 1) // 3, R) * pow(wi, j, R) % R])
        assert (got[j] == zo.poly_eval(coeffs, zo.x_table(x, log_n))[0]).all(), j
    # to_coset = 1 scales the whole extended buffer in place
    call("extended_prepare", dv.ptr(tmp if swap.value else s), dv.ptr(dcp), 3, n, N, 1,
         dv.stream_handle())
    got2 = dv.from_dev(tmp if swap.value else s)
    assert (got2 == zo.extended_prepare(got, cp, n, N, 1)).all()



def _ntt_batch(zo, specs):
    """zk_ntt_batch over [(x, log_n, omega, n_inv or None)]; returns outputs"""
    import ctypes
    from zkprover import device as dv
    from zkprover._lib import call, lib

    class Item(ctypes.Structure):
        _fields_ = [("buf", ctypes.c_void_p), ("tmp", ctypes.c_void_p), ("omegas", ctypes.c_void_p),
                    ("n_inv", ctypes.c_void_p), ("log_n", ctypes.c_int), ("reserved", ctypes.c_int)]
    keep, bufs = [], []
    items = (Item * len(specs))()
    for m, (x, log_n, omega, n_inv) in enumerate(specs):
        om = _omegas_dev(zo, log_n, omega)
        b, t = dv.to_dev(x), dv.empty(len(x))
        ni = dv.to_dev(n_inv) if n_inv is not None else None
        keep += [om, b, t, ni]
        bufs.append((b, t))
        items[m] = Item(dv.ptr(b), dv.ptr(t), dv.ptr(om), dv.ptr(ni) if ni is not None else None, log_n, 0)
    swapped = (ctypes.c_ubyte * len(specs))()
    lib().zk_ntt_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    call("zk_ntt_batch", ctypes.addressof(items), len(specs), ctypes.addressof(swapped), dv.stream_handle())
    return [dv.from_dev(t if swapped[m] else b) for m, (b, t) in enumerate(bufs)]


def test_ntt_batch_mixed_sizes(gpu, zo):
    """NTTs and iNTTs of sizes 2^0..2^17 in one batch (one launch per pass
    round); each equals the oracle's transform of its input"""
    logs = [0, 1, 3, 9, 10, 11, 14, 16, 17, 12, 5]
    specs, exps = [], []
    for k, log_n in enumerate(logs):
        n = 1 << log_n
        x = zo.rand_fr(n, 300 + k)
        w = zo.root_of_unity(log_n)
        if k % 2:
            w_inv = zo.fr_inv(w)
            n_inv = zo.fr_inv(zo.fr_from_u64(n))
            specs.append((x, log_n, w_inv, n_inv))
            exps.append(zo.ntt(x, log_n, w_inv, n_inv))
        else:
            specs.append((x, log_n, w, None))
            exps.append(zo.ntt(x, log_n, w))
    got = _ntt_batch(zo, specs)
    for k in range(len(logs)):
        assert (got[k] == exps[k]).all(), logs[k]


def test_ntt_batch_many_items(gpu, zo):
    """more transforms than one launch carries (48): the pass rounds split
    over several launches"""
    specs, exps = [], []
    for k in range(100):
        log_n = 6 + (k % 7)
        x = zo.rand_fr(1 << log_n, 500 + k)
        w = zo.root_of_unity(log_n)
        specs.append((x, log_n, w, None))
        exps.append(zo.ntt(x, log_n, w))
    got = _ntt_batch(zo, specs)
    for k in range(100):
        assert (got[k] == exps[k]).all(), k


@pytest.mark.parametrize("max_deg", [1, 3, 5, 6, 7, 8])
def test_ntt_radix_choice_same_result(gpu, zo, max_deg):
    """every radix plan the autotuner may pick gives the same transform"""
    import ctypes
    from zkprover._lib import lib
    log_n = 13
    x = zo.rand_fr(1 << log_n, 77)
    w = zo.root_of_unity(log_n)
    L = lib()
    L.zk_ntt_set_max_deg.argtypes = [ctypes.c_int, ctypes.c_int]
    assert L.zk_ntt_set_max_deg(log_n, max_deg) == 0
    try:
        got = _gpu_ntt(zo, x, log_n, w)
    finally:
        L.zk_ntt_set_max_deg(log_n, 0)
    assert (got == zo.ntt(x, log_n, w)).all()


@pytest.mark.parametrize("op", [0, 1, 2, 3, 4])
def test_fp29_ops_match_oracle(gpu, zo, op):
    """the 29-bit-limb lazy Montgomery field of the bucket accumulation, through
    its conversions from and to the 32-bit form, equals the oracle's Fp
    arithmetic (edge values 0, 1, p-1, p-2 and 2^253 included)"""
    from zkprover import device as dv
    from zkprover._lib import call
    rng = np.random.default_rng(29 + op)
    xs = [0, 1, P - 1, P - 2, 1 << 253, P // 2] + [int.from_bytes(rng.bytes(32), "little") % P for _ in range(4090)]
    ys = xs[::-1]
    a = np.concatenate([zo.fp_to_mont_int(x) for x in xs])
    b = np.concatenate([zo.fp_to_mont_int(y) for y in ys])
    n = len(xs)
    da, db, do = dv.to_dev(a), dv.to_dev(b), dv.empty(n)
    call("zk_fp29_op", dv.ptr(do), dv.ptr(da), dv.ptr(db), op, n, dv.stream_handle())
    got = dv.from_dev(do)
    exp = {0: [x * y % P for x, y in zip(xs, ys)], 1: [(x + y) % P for x, y in zip(xs, ys)],
           2: [(x - y) % P for x, y in zip(xs, ys)], 3: [x * x % P for x in xs], 4: xs}[op]
    assert [zo.fp_from_mont_int(got[i]) for i in range(n)] == exp


class _CosetItem(ctypes.Structure):
    _fields_ = [("src", ctypes.c_void_p), ("buf", ctypes.c_void_p), ("tmp", ctypes.c_void_p),
                ("omegas", ctypes.c_void_p), ("log_n", ctypes.c_int), ("log_ext", ctypes.c_int)]


def _coset_ref(zo, dv, call, coeffs, cp, log_n, log_ext, om):
    """the three-call sequence: copy into an extended buffer, extended_prepare, ntt"""
    import ctypes
    n, N = 1 << log_n, 1 << (log_n + log_ext)
    host = np.zeros((N, 4), dtype=np.uint64)
    host[:n] = coeffs
    s, tmp = dv.to_dev(host), dv.empty(N)
    dcp = dv.to_dev(cp)
    call("extended_prepare", dv.ptr(s), dv.ptr(dcp), len(cp) + 1, n, N, 0, dv.stream_handle())
    swap = ctypes.c_bool(False)
    call("ntt", dv.ptr(s), dv.ptr(tmp), None, dv.ptr(om), None, log_n + log_ext, 8, ctypes.addressof(swap),
         dv.stream_handle())
    return dv.from_dev(tmp if swap.value else s)


@pytest.mark.parametrize("shapes", [[(10, 2)], [(22, 2)], [(12, 1), (9, 3), (16, 2), (4, 2)]])
def test_coset_extend_batch_equals_three_calls(gpu, zo, shapes):
    import ctypes
    from zkprover import device as dv
    from zkprover._lib import call, lib
    f = lib().zk_coset_extend_batch
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    f.restype = ctypes.c_int
    zeta = zo.to_mont([pow(7, (R - 1) // 3, R)])
    cp = np.concatenate([zeta, zo.fr_mul(zeta, zeta)])
    keep, items, exps = [], (_CosetItem * len(shapes))(), []
    for m, (log_n, log_ext) in enumerate(shapes):
        n, N = 1 << log_n, 1 << (log_n + log_ext)
        coeffs = zo.rand_fr(n, 900 + 7 * log_n + log_ext)
        om = _omegas_dev(zo, log_n + log_ext, zo.root_of_unity(log_n + log_ext))
        exps.append(_coset_ref(zo, dv, call, coeffs, cp, log_n, log_ext, om))
        src, buf, tmp = dv.to_dev(coeffs), dv.empty(N), dv.empty(N)
        keep += [src, buf, tmp, om]
        items[m] = _CosetItem(src.data_ptr(), buf.data_ptr(), tmp.data_ptr(), om.data_ptr(), log_n, log_ext)
    dcp = dv.to_dev(cp)
    sw = (ctypes.c_ubyte * len(shapes))()
    assert f(ctypes.addressof(items), len(shapes), dcp.data_ptr(), 3, sw, dv.stream_handle()) == 0
    for m in range(len(shapes)):
        got = dv.from_dev(keep[4 * m + 2] if sw[m] else keep[4 * m + 1])
        assert (got == exps[m]).all(), shapes[m]
    # a source that aliases the output is refused
    bad = (_CosetItem * 1)(_CosetItem(keep[1].data_ptr(), keep[1].data_ptr(), keep[2].data_ptr(), keep[3].data_ptr(),
                                      shapes[0][0], shapes[0][1]))
    assert f(ctypes.addressof(bad), 1, dcp.data_ptr(), 3, sw, dv.stream_handle()) != 0
