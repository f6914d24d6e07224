"""GPU parity: multi-open kernels through the C-ABI.

poly_eval (src/cuda/bn254_c.rs:200), shplonk_h_x_merge (:208) and
shplonk_h_x_div_points (:218) are compared bit-for-bit with the oracle's
restatements of cuda/bn254.cu:238-256, :1133-1170, :1608-1633, and
zk_kate_divide with the oracle's restatement of the host division by (X - z)
(src/multiopen.rs:145-159, :543-553).  Sizes straddle the 2048-element scan
block and, at 2^23 + 3, the 2048-block chunk of the top-level scan.  At the
largest sizes the division is also checked through q(t)(t - z) + a(z) = a(t).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

R = 21888242871839275222246405745257275088548364400416034343698204186575808495617


def _dv():
    from zkprover import device as dv
    from zkprover._lib import call, lib
    return dv, call, lib


def _horner_ints(zo, p, x):
    acc = 0
    xi = zo.from_mont(x)[0]
    for c in reversed(zo.from_mont(p)):
        acc = (acc * xi + c) % R
    return acc


def _log2_ceil(n):
    return max(1, (n - 1).bit_length())


def _gpu_poly_eval(p, xt):
    dv, call, _ = _dv()
    n = len(p)
    dp, dx = dv.to_dev(p), dv.to_dev(xt)
    res = dv.empty(n)
    call("poly_eval", dv.ptr(dp), dv.ptr(res), dv.ptr(dx), n, dv.stream_handle())
    return dv.from_dev(res[:1])


@pytest.mark.parametrize("log_n", [1, 4, 5, 11, 12, 16, 20, 22])
def test_poly_eval_matches_oracle(gpu, zo, log_n):
    n = 1 << log_n
    p = zo.rand_fr(n, 100 + log_n)
    xt = zo.x_table(zo.rand_fr(1, 200 + log_n), log_n)
    assert (_gpu_poly_eval(p, xt) == zo.poly_eval(p, xt)).all()


@pytest.mark.parametrize("n", [1, 3, 2047, 2049, 4097, 5000])
def test_poly_eval_any_length(gpu, zo, n):
    """lengths the reference does not take (not a power of two): Horner"""
    p = zo.rand_fr(n, n)
    x = zo.rand_fr(1, n + 1)
    xt = zo.x_table(x, _log2_ceil(n))
    assert zo.from_mont(_gpu_poly_eval(p, xt))[0] == _horner_ints(zo, p, x)


def _gpu_poly_eval_batch(polys, xt):
    """zk_poly_eval_batch over equal-length polynomials; returns (count, 4)"""
    dv, call, lib = _dv()
    n, count = len(polys[0]), len(polys)
    dps = [dv.to_dev(p) for p in polys]
    dx = dv.to_dev(xt)
    out = dv.zeros(count)
    L = lib()
    L.zk_poly_eval_batch_scratch_bytes.restype = ctypes.c_size_t
    L.zk_poly_eval_batch_scratch_bytes.argtypes = [ctypes.c_int, ctypes.c_int]
    L.zk_poly_eval_batch.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                     ctypes.c_void_p, ctypes.c_void_p]
    nbytes = L.zk_poly_eval_batch_scratch_bytes(n, count)
    import torch
    scratch = torch.empty(nbytes, dtype=torch.uint8, device=dv.device())
    pa = (ctypes.c_void_p * count)(*[d.data_ptr() for d in dps])
    oa = (ctypes.c_void_p * count)(*[out.data_ptr() + 32 * m for m in range(count)])
    call("zk_poly_eval_batch", ctypes.addressof(pa), ctypes.addressof(oa), count, dv.ptr(dx), n, scratch.data_ptr(),
         dv.stream_handle())
    return dv.from_dev(out)


@pytest.mark.parametrize("n,count", [(1, 3), (3, 2), (2047, 3), (2049, 3), (5000, 2), (4096, 70), (1 << 16, 4),
                                     ((1 << 23) + (1 << 12) + 5, 2)])
def test_poly_eval_batch_matches_oracle(gpu, zo, n, count):
    """many polynomials at one point in one call (more than the 64 of one
    launch at count 70; two chunks of block values at 2^23 + 2^12 + 5), each
    equal to the oracle's poly_eval"""
    polys = [zo.rand_fr(n, 7000 + 31 * m + n % 1000) for m in range(count)]
    lg = _log2_ceil(n)
    xt = zo.x_table(zo.rand_fr(1, 300 + count), lg)
    got = _gpu_poly_eval_batch(polys, xt)
    for m in range(count):
        # the oracle takes power-of-two lengths: zero coefficients up to 2^lg
        padded = zo.fr_zeros(1 << lg)
        padded[:n] = polys[m]
        assert (got[m:m + 1] == zo.poly_eval(padded, xt)).all(), (n, m)


def test_poly_eval_special_points(gpu, zo):
    n = 1 << 13
    p = zo.rand_fr(n, 1)
    assert (_gpu_poly_eval(p, zo.x_table(zo.fr_zeros(1), 13)) == p[:1]).all()
    one = zo.to_mont([1])
    assert zo.from_mont(_gpu_poly_eval(p, zo.x_table(one, 13)))[0] == sum(zo.from_mont(p)) % R


def test_poly_eval_two_chunks(gpu, zo):
    """2^23 + 2^12 + 5 coefficients: 4099 scan blocks, two top-level chunks"""
    n = (1 << 23) + (1 << 12) + 5
    p = zo.rand_fr(n, 7)
    x = zo.rand_fr(1, 8)
    xt = zo.x_table(x, 24)
    got = _gpu_poly_eval(p, xt)
    pad = np.concatenate([p, zo.fr_zeros((1 << 24) - n)])
    assert (got == zo.poly_eval(pad, xt)).all()


def _kate(dv, call, lib, a, z, in_place, want_eval=True):
    n = len(a)
    lib().zk_kate_divide_scratch_bytes.restype = ctypes.c_size_t
    lib().zk_kate_divide_scratch_bytes.argtypes = [ctypes.c_int]
    tb = lib().zk_kate_divide_scratch_bytes(n)
    temp = dv.empty((tb + 31) // 32)
    din = dv.to_dev(a)
    dout = din if in_place else dv.empty(n)
    dz = dv.to_dev(z)
    ev = dv.zeros(1) if want_eval else None
    call("zk_kate_divide", dv.ptr(dout), dv.ptr(din), dv.ptr(dz), dv.ptr(ev), n, dv.ptr(temp),
         dv.stream_handle())
    return dv.from_dev(dout), (dv.from_dev(ev) if want_eval else None)


@pytest.mark.parametrize("n,in_place", [(1, True), (2, False), (3, True), (2047, False), (2048, True),
                                        (2049, True), (4097, False), (1 << 16, True), (1 << 20, False)])
def test_kate_divide_matches_oracle(gpu, zo, n, in_place):
    dv, call, lib = _dv()
    a = zo.rand_fr(n, 300 + n)
    z = zo.rand_fr(1, 400 + n)
    got, ev = _kate(dv, call, lib, a, z, in_place)
    assert (got == zo.kate_divide(a, z)).all()
    if n > 1:
        xt = zo.x_table(z, _log2_ceil(n))
        pad = np.concatenate([a, zo.fr_zeros((1 << _log2_ceil(n)) - n)])
        assert (ev == zo.poly_eval(pad, xt)).all()


def test_kate_divide_identity_large(gpu, zo):
    """q(t) (t - z) + a(z) = a(t) at a random t, n = 2^23 + 3 (two top chunks)"""
    dv, call, lib = _dv()
    n = (1 << 23) + 3
    a = zo.rand_fr(n, 11)
    z = zo.rand_fr(1, 12)
    q, az = _kate(dv, call, lib, a, z, True)
    assert not q[n - 1].any()
    t = zo.rand_fr(1, 13)
    xt = zo.x_table(t, 24)
    pad = lambda v: np.concatenate([v, zo.fr_zeros((1 << 24) - n)])
    at = zo.poly_eval(pad(a), xt)
    qt = zo.poly_eval(pad(q), xt)
    assert (zo.fr_add(zo.fr_mul(qt, zo.fr_sub(t, z)), az) == at).all()


@pytest.mark.parametrize("n,diff_n", [(1, 0), (64, 1), (1000, 3), (1 << 16, 2)])
def test_h_x_merge_matches_oracle(gpu, zo, n, diff_n):
    dv, call, _ = _dv()
    res, values, omegas = zo.rand_fr(n, 1), zo.rand_fr(n, 2), zo.rand_fr(n, 3)
    v = zo.rand_fr(1, 4)
    diff = zo.rand_fr(max(diff_n, 1), 5)[:diff_n]
    exp = zo.shplonk_h_x_merge(res, v, values, omegas, diff)
    d_res, d_v, d_val, d_om = dv.to_dev(res), dv.to_dev(v), dv.to_dev(values), dv.to_dev(omegas)
    d_diff = dv.to_dev(diff) if diff_n else None
    call("shplonk_h_x_merge", dv.ptr(d_res), dv.ptr(d_v), dv.ptr(d_val), dv.ptr(d_om), dv.ptr(d_diff), diff_n, n)
    import torch
    torch.cuda.synchronize()
    assert (dv.from_dev(d_res) == exp).all()
    # stream form, applied again on the result
    exp2 = zo.shplonk_h_x_merge(exp, v, values, omegas, diff)
    call("zk_shplonk_h_x_merge_on", dv.ptr(d_res), dv.ptr(d_v), dv.ptr(d_val), dv.ptr(d_om), dv.ptr(d_diff),
         diff_n, n, dv.stream_handle())
    assert (dv.from_dev(d_res) == exp2).all()


@pytest.mark.parametrize("n,points_n", [(1, 1), (100, 2), (1024, 1), (1025, 3), (1 << 16, 4), (1 << 20, 3)])
def test_h_x_div_points_matches_oracle(gpu, zo, n, points_n):
    dv, call, _ = _dv()
    values = zo.rand_fr(n, 21)
    log_n = _log2_ceil(n)
    omegas = zo.expand_omega(1 << log_n, zo.root_of_unity(log_n))[:n]
    points = zo.rand_fr(points_n, 22)
    exp = zo.shplonk_h_x_div_points(values, omegas, points)
    d_val, d_om, d_pt = dv.to_dev(values), dv.to_dev(omegas), dv.to_dev(points)
    call("shplonk_h_x_div_points", dv.ptr(d_val), dv.ptr(d_om), dv.ptr(d_pt), points_n, n)
    import torch
    torch.cuda.synchronize()
    assert (dv.from_dev(d_val) == exp).all()


def test_h_x_div_points_zero_denominator(gpu, zo):
    dv, call, _ = _dv()
    log_n = 12
    n = 1 << log_n
    values = zo.rand_fr(n, 31)
    omegas = zo.expand_omega(n, zo.root_of_unity(log_n))
    points = np.concatenate([zo.rand_fr(1, 32), omegas[777:778]])
    exp = zo.shplonk_h_x_div_points(values, omegas, points)
    d_val, d_om, d_pt = dv.to_dev(values), dv.to_dev(omegas), dv.to_dev(points)  # held until the kernel ran
    call("zk_shplonk_h_x_div_points_on", dv.ptr(d_val), dv.ptr(d_om), dv.ptr(d_pt), 2, n, dv.stream_handle())
    got = dv.from_dev(d_val)
    assert not got[777].any()
    assert (got == exp).all()


def test_multiopen_invalid_arguments(gpu, zo):
    dv, call, lib = _dv()
    buf = dv.zeros(4)
    p = dv.ptr(buf)
    s = dv.stream_handle()
    hipErrorInvalidValue = 1
    assert lib().poly_eval(p, p, p, 0, s) == hipErrorInvalidValue
    assert lib().poly_eval(None, p, p, 4, s) == hipErrorInvalidValue
    assert lib().zk_kate_divide(p, p, p, None, 0, p, s) == hipErrorInvalidValue
    assert lib().zk_kate_divide(p, p, None, None, 4, p, s) == hipErrorInvalidValue
    assert lib().shplonk_h_x_div_points(p, p, p, 0, 4) == hipErrorInvalidValue
    assert lib().shplonk_h_x_merge(p, p, p, p, None, 2, 4) == hipErrorInvalidValue
