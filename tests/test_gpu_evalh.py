"""GPU parity: eval_h evaluator and constraint kernels (cuda/bn254.cu:542-1213).

Each kernel is run on seeded random operands at extended-domain sizes with
the rotations eval_h uses (+-2^(k_ext - k)), and compared bit-for-bit with
the oracle's sequential restatement of the same per-row formula and fold
order.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _env():
    from zkprover import device as dv
    from zkprover._lib import call
    from zkprover.device import Arena
    return dv, call, Arena()


def _r(zo, n, seed):
    return zo.rand_fr(n, seed)


@pytest.mark.parametrize("op", [0, 1, 2, 3])
@pytest.mark.parametrize("rots", [(0, 0), (4, -4), (-1, 3)])
def test_field_op(gpu, zo, op, rots):
    dv, call, A = _env()
    n = 1 << 12
    l, r = _r(zo, n, 1), _r(zo, n, 2)
    lc, rc = _r(zo, 1, 3), _r(zo, 1, 4)
    res = dv.empty(n)
    for use_lc, use_rc, use_r in [(False, False, True), (True, True, True), (False, True, False)]:
        call("field_op", dv.ptr(res), A.put(l), rots[0], A.put(lc) if use_lc else None,
             A.put(r) if use_r else None, rots[1], A.put(rc) if use_rc else None, n, op, dv.stream_handle())
        exp = zo.field_op(l, rots[0], lc if use_lc else None, r if use_r else None, rots[1],
                          rc if use_rc else None, n, op)
        assert (dv.from_dev(res) == exp).all()


@pytest.mark.parametrize("log_n", [10, 16])
def test_field_op_batch_mul_sum(gpu, zo, log_n):
    dv, call, A = _env()
    n = 1 << log_n
    cols = [_r(zo, n, 10 + j) for j in range(6)]
    dcols = [dv.to_dev(c) for c in cols]
    for t in dcols:
        A.hold(t)
    rng = np.random.default_rng(log_n)
    terms = []
    for t in range(9):
        k = int(rng.integers(0, 4))  # 0..3 factors (0: the constant term)
        coeff = _r(zo, 1, 100 + t)
        facs = [(int(rng.integers(0, 6)), int(rng.choice([0, 4, -4, 8, -1]))) for _ in range(k)]
        terms.append((coeff, facs))
    # device term list [c0, a00, a01, NULL, c1, ...] and rotation list
    import torch
    coeff_t = [dv.to_dev(c) for c, _ in terms]
    for t in coeff_t:
        A.hold(t)
    ptrs, rots = [], []
    for (c, facs), ct in zip(terms, coeff_t):
        ptrs.append(ct.data_ptr())
        for j, rot in facs:
            ptrs.append(dcols[j].data_ptr())
            rots.append(rot)
        ptrs.append(0)
    dptrs = torch.from_numpy(np.array(ptrs, dtype=np.uint64).view(np.int64)).to(dv.device())
    drots = torch.from_numpy(np.array(rots + [0], dtype=np.int32)).to(dv.device())
    A.hold(dptrs)
    A.hold(drots)
    res0 = _r(zo, n, 99)
    res = dv.to_dev(res0)
    call("field_op_batch_mul_sum", dv.ptr(res), dptrs.data_ptr(), drots.data_ptr(), len(ptrs), n,
         dv.stream_handle())
    exp = zo.field_op_batch_mul_sum(res0, [(c, [(cols[j], rot) for j, rot in facs]) for c, facs in terms], n)
    assert (dv.from_dev(res) == exp).all()


def test_field_mul_zip(gpu, zo):
    dv, call, A = _env()
    n = 1 << 14
    buf, coeff = _r(zo, n, 1), _r(zo, 12, 2)
    d = dv.to_dev(buf)
    call("field_mul_zip", dv.ptr(d), A.put(coeff), 12, n)
    assert (dv.from_dev(d) == zo.field_mul_zip(buf, coeff)).all()


@pytest.mark.parametrize("log_n", [12, 16])
def test_permutation_kernels(gpu, zo, log_n):
    dv, call, A = _env()
    n = 1 << log_n
    res0 = _r(zo, n, 1)
    first, last, l0, ll, y = _r(zo, n, 2), _r(zo, n, 3), _r(zo, n, 4), _r(zo, n, 5), _r(zo, 1, 6)
    d = dv.to_dev(res0)
    call("permutation_eval_h_p1", dv.ptr(d), A.put(first), A.put(last), A.put(l0), A.put(ll), A.put(y), n)
    exp = zo.permutation_eval_h_p1(res0, first, last, l0, ll, y)
    assert (dv.from_dev(d) == exp).all()
    sets = [_r(zo, n, 20 + j) for j in range(4)]
    rot = -(1 << 2) * 5
    call("permutation_eval_h_p2", dv.ptr(d), A.put_all(sets), A.put(l0), A.put(ll), A.put(y), 4, rot, n)
    exp = zo.permutation_eval_h_p2(exp, sets, l0, y, rot)
    assert (dv.from_dev(d) == exp).all()
    call("logup_eval_h_z_set", dv.ptr(d), A.put_all(sets[:3]), A.put(l0), A.put(ll), A.put(y), 3, rot, n,
         dv.stream_handle())
    exp = zo.permutation_eval_h_p2(exp, sets[:3], l0, y, rot)
    assert (dv.from_dev(d) == exp).all()
    beta, gamma, p = _r(zo, 1, 30), _r(zo, 1, 31), _r(zo, n, 32)
    call("permutation_eval_h_l", dv.ptr(d), A.put(beta), A.put(gamma), A.put(p), n, dv.stream_handle())
    exp = zo.permutation_eval_h_l(exp, beta, gamma, p)
    assert (dv.from_dev(d) == exp).all()


@pytest.mark.parametrize("log_n", [12, 16])
def test_logup_kernels(gpu, zo, log_n):
    dv, call, A = _env()
    n = 1 << log_n
    rot = 4
    a = [_r(zo, n, 40 + j) for j in range(9)]
    y = _r(zo, 1, 60)
    res0 = _r(zo, n, 61)
    d = dv.to_dev(res0)
    call("logup_eval_h", dv.ptr(d), *[A.put(x) for x in a], A.put(y), rot, n, dv.stream_handle())
    exp = zo.logup_eval_h(res0, *a, y, rot)
    assert (dv.from_dev(d) == exp).all()
    call("logup_eval_h_extra_inputs", dv.ptr(d), A.put(a[0]), A.put(a[1]), A.put(a[4]), A.put(a[8]), A.put(y),
         rot, n, dv.stream_handle())
    exp = zo.logup_eval_h_extra_inputs(exp, a[0], a[1], a[4], a[8], y, rot)
    assert (dv.from_dev(d) == exp).all()
    sets = [_r(zo, n, 70 + j) for j in range(4)]
    prod, psum = dv.empty(n), dv.empty(n)
    call("logup_eval_h_inputs_product_sum", dv.ptr(prod), dv.ptr(psum), A.put_all(sets), 4, n, dv.stream_handle())
    ep, es = zo.logup_eval_h_inputs_product_sum(sets)
    assert (dv.from_dev(prod) == ep).all()
    assert (dv.from_dev(psum) == es).all()


@pytest.mark.parametrize("log_n", [12, 16])
def test_shuffle_kernels(gpu, zo, log_n):
    dv, call, A = _env()
    n = 1 << log_n
    rot = 4
    inp, tab, z, l0, ll, la = [_r(zo, n, 80 + j) for j in range(6)]
    y = _r(zo, 1, 90)
    res0 = _r(zo, n, 91)
    d = dv.to_dev(res0)
    call("shuffle_eval_h", dv.ptr(d), A.put(inp), A.put(tab), A.put(z), A.put(l0), A.put(ll), A.put(la), A.put(y),
         rot, n)
    exp = zo.shuffle_eval_h(res0, inp, tab, z, l0, ll, la, y, rot)
    assert (dv.from_dev(d) == exp).all()
    g = 3
    ins = [_r(zo, n, 100 + j) for j in range(g)]
    tbs = [_r(zo, n, 110 + j) for j in range(g)]
    betas = _r(zo, g, 120)
    call("shuffle_eval_h_v2", dv.ptr(d), A.put_all(ins), A.put_all(tbs), A.put(betas), g, A.put(z), A.put(l0),
         A.put(ll), A.put(la), A.put(y), rot, n)
    exp = zo.shuffle_eval_h_v2(exp, ins, tbs, betas, z, l0, ll, la, y, rot)
    assert (dv.from_dev(d) == exp).all()


@pytest.mark.parametrize("g,rot", [(1, 1), (1, 4), (16, 4)])
def test_shuffle_v2_group_counts(gpu, zo, g, rot):
    """one group (the prover core's call) and 16 groups (the size of the
    reference's betas vector, src/eval_h.rs:1080-1083), bit-exact"""
    dv, call, A = _env()
    n = 1 << 12
    z, l0, ll, la = [_r(zo, n, 130 + j) for j in range(4)]
    y, res0 = _r(zo, 1, 135), _r(zo, n, 136)
    ins = [_r(zo, n, 140 + j) for j in range(g)]
    tbs = [_r(zo, n, 160 + j) for j in range(g)]
    betas = _r(zo, g, 180)
    d = dv.to_dev(res0)
    call("shuffle_eval_h_v2", dv.ptr(d), A.put_all(ins), A.put_all(tbs), A.put(betas), g, A.put(z), A.put(l0),
         A.put(ll), A.put(la), A.put(y), rot, n)
    exp = zo.shuffle_eval_h_v2(res0, ins, tbs, betas, z, l0, ll, la, y, rot)
    assert (dv.from_dev(d) == exp).all()


@pytest.mark.parametrize("nc,rot", [(1, 4), (2, 4), (3, -4)])
def test_permutation_eval_h_set(gpu, zo, nc, rot):
    """the fused permutation-set term equals the same expression built from the
    oracle's element-wise field operations (the reference's field_op sequence,
    src/eval_h.rs:700-840)"""
    dv, call, A = _env()
    n = 1 << 12
    h, z, X, la = _r(zo, n, 40), _r(zo, n, 41), _r(zo, n, 42), _r(zo, n, 43)
    beta, gamma, y = _r(zo, 1, 44), _r(zo, 1, 45), _r(zo, 1, 46)
    vs = [_r(zo, n, 50 + c) for c in range(nc)]
    ss = [_r(zo, n, 60 + c) for c in range(nc)]
    bd = _r(zo, nc, 70)
    # l = z(wX) prod (v + beta sigma + gamma), r = z prod (v + gamma + bd X)
    l = zo.field_op(z, rot, None, None, 0, None, n, 2)
    r = np.array(z, copy=True)
    for c in range(nc):
        vg = zo.field_op(vs[c], 0, None, None, 0, gamma, n, 0)
        l = zo.field_op(l, 0, None, zo.field_op(ss[c], 0, beta, vg, 0, None, n, 0), 0, None, n, 1)
        r = zo.field_op(r, 0, None, zo.field_op(X, 0, bd[c:c + 1], vg, 0, None, n, 0), 0, None, n, 1)
    t = zo.field_op(zo.field_op(l, 0, None, r, 0, None, n, 3), 0, None, la, 0, None, n, 1)
    exp = zo.field_op(h, 0, y, t, 0, None, n, 0)
    th = dv.to_dev(h)
    call("zk_permutation_eval_h_set", A.hold(th), A.put(z), A.put_all(vs), A.put_all(ss), nc, A.put(beta),
         A.put(gamma), A.put(bd), A.put(X), A.put(la), A.put(y), rot, n, dv.stream_handle())
    assert (dv.from_dev(th) == exp).all()


@pytest.mark.parametrize("log_n", [12, 16])
def test_field_sum(gpu, zo, log_n):
    """bn254_c.rs:71 field_sum: rotated coefficient sums, a null coefficient
    pointer, negative and wrapping rotations, and no terms at all"""
    import torch
    dv, call, A = _env()
    n = 1 << log_n
    omegas = zo.expand_omega(n, zo.root_of_unity(log_n))
    vs = [_r(zo, n, 200 + j) for j in range(5)]
    cs = [None, _r(zo, 1, 210), None, _r(zo, 1, 211), _r(zo, 1, 212)]
    rots = [0, 4, -4, n + 3, -(n - 1)]
    vptr = A.put_all(vs)
    cptr_host = np.array([0 if c is None else A.put(c) for c in cs], dtype=np.uint64)
    cptr = A.hold(torch.from_numpy(cptr_host.view(np.int64)).to(dv.device()))
    rptr = A.hold(torch.tensor(rots, dtype=torch.int32, device=dv.device()))
    res = dv.empty(n)
    call("field_sum", dv.ptr(res), vptr, cptr, rptr, A.put(omegas), len(vs), n)
    torch.cuda.synchronize()
    assert (dv.from_dev(res) == zo.field_sum(vs, cs, rots, omegas)).all()
    call("field_sum", dv.ptr(res), vptr, cptr, rptr, A.put(omegas), 0, n)
    torch.cuda.synchronize()
    assert (dv.from_dev(res) == 0).all()


@pytest.mark.parametrize("log_n,rot", [(12, 4), (16, 4), (12, 1)])
def test_lookup_eval_h(gpu, zo, log_n, rot):
    """bn254_c.rs:90 lookup_eval_h against the oracle's five folds"""
    dv, call, A = _env()
    n = 1 << log_n
    inp, tab, pin, ptab, z, l0, ll, la = [_r(zo, n, 220 + j) for j in range(8)]
    y, beta, gamma = _r(zo, 1, 230), _r(zo, 1, 231), _r(zo, 1, 232)
    res0 = _r(zo, n, 233)
    d = dv.to_dev(res0)
    call("lookup_eval_h", dv.ptr(d), *[A.put(x) for x in (inp, tab, pin, ptab, z, l0, ll, la, y, beta, gamma)],
         rot, n, dv.stream_handle())
    exp = zo.lookup_eval_h(res0, inp, tab, pin, ptab, z, l0, ll, la, y, beta, gamma, rot)
    assert (dv.from_dev(d) == exp).all()
