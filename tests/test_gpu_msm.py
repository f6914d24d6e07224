"""GPU parity: BN254 G1 MSM (msm + batch_msm_collect drop-in, zk_msm, zk_msm_batch).

Pinning: the CPU oracle's MSM is itself checked against the naive sum of
scalar multiples and against the reference's known-answer construction
(src/cuda/msm.rs:457-537: bases c_{i mod 256} * G, expected
(sum_i s_i c_{i mod 256}) * G, a closed form independent of any MSM code).
GPU results are compared as affine points (canonical bytes).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

R = 21888242871839275222246405745257275088548364400416034343698204186575808495617

OUT_BYTES = 64 * 1024


def _zk_msm(points, scalars, c=0):
    import torch
    from zkprover import device as dv
    from zkprover._lib import call, lib
    n = len(scalars)
    need = lib().zk_msm_scratch_bytes
    need.restype = ctypes.c_size_t
    need.argtypes = [ctypes.c_int, ctypes.c_int]
    nbytes = need(n, c)
    scratch = torch.empty(nbytes, dtype=torch.uint8, device=dv.device())
    out = torch.zeros(OUT_BYTES, dtype=torch.uint8, device=dv.device())
    dp = dv.to_dev(points.reshape(-1, 4))
    ds = dv.to_dev(scalars)
    lib().zk_msm.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                                      ctypes.c_size_t, ctypes.c_void_p]
    call("zk_msm", out.data_ptr(), dp.data_ptr(), ds.data_ptr(), n, c, scratch.data_ptr(), nbytes,
         dv.stream_handle())
    jac = out[:96].cpu().numpy().view(np.uint64).reshape(1, 12).copy()
    return jac


class _Item(ctypes.Structure):
    _fields_ = [("points", ctypes.c_void_p), ("scalars", ctypes.c_void_p), ("n", ctypes.c_int),
                ("reserved", ctypes.c_int)]


def _zk_msm_batch(point_sets, scalar_sets, c=0):
    """zk_msm_batch over (points, scalars) pairs; returns (count, 12) Jacobian"""
    import torch
    from zkprover import device as dv
    from zkprover._lib import call, lib
    count = len(scalar_sets)
    keep = []
    items = (_Item * count)()
    for m, (p, s) in enumerate(zip(point_sets, scalar_sets)):
        dp, ds = dv.to_dev(p.reshape(-1, 4)), dv.to_dev(s)
        keep += [dp, ds]
        items[m] = _Item(dp.data_ptr(), ds.data_ptr(), len(s), 0)
    need = lib().zk_msm_batch_scratch_bytes
    need.restype = ctypes.c_size_t
    need.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    nbytes = need(ctypes.addressof(items), count, c)
    assert nbytes > 0
    scratch = torch.empty(nbytes, dtype=torch.uint8, device=dv.device())
    out = torch.zeros(count * OUT_BYTES, dtype=torch.uint8, device=dv.device())
    lib().zk_msm_batch.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                   ctypes.c_size_t, ctypes.c_void_p]
    call("zk_msm_batch", out.data_ptr(), ctypes.addressof(items), count, c, scratch.data_ptr(), nbytes,
         dv.stream_handle())
    o = out.view(count, OUT_BYTES)[:, :96].cpu().numpy().copy()
    return o.view(np.uint64).reshape(count, 12)


def _affine_of(zo, jac):
    return zo.g1_to_affine(jac)


def _batch_msm(points, scalar_sets, skip_zero=0):
    """the reference host driver's path (src/cuda/msm.rs:239-455) via zkp_batch_msm"""
    from zkprover import device as dv
    from zkprover._lib import lib, check
    dp = dv.to_dev(points.reshape(-1, 4))
    ds = [dv.to_dev(s) for s in scalar_sets]
    arr = (ctypes.c_void_p * len(ds))(*[d.data_ptr() for d in ds])
    out = np.zeros((len(ds), 8), dtype=np.uint64)
    f = lib().zkp_batch_msm
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    f.restype = ctypes.c_int
    check(f(out.ctypes.data, dp.data_ptr(), ctypes.cast(arr, ctypes.c_void_p), len(ds),
            len(scalar_sets[0]), skip_zero), "zkp_batch_msm")
    return out


@pytest.mark.parametrize("n", [1, 2, 3, 5, 64, 1000, 4096])
def test_zk_msm_matches_oracle(gpu, zo, n):
    pts = zo.gen_bases_random(n, 1000 + n)
    sc = zo.rand_fr(n, 2000 + n)
    got = _affine_of(zo, _zk_msm(pts, sc))
    exp = zo.g1_to_affine(zo.msm(pts, sc))
    assert (got == exp).all()
    if n <= 64:
        assert (exp == zo.g1_to_affine(zo.msm_naive(pts, sc))).all()


@pytest.mark.parametrize("c", [3, 5, 8, 11, 13, 16])
def test_zk_msm_window_sizes(gpu, zo, c):
    n = 1 << 13
    pts = zo.gen_bases_ap(n, 31)
    sc = zo.rand_fr(n, 32)
    got = _affine_of(zo, _zk_msm(pts, sc, c))
    assert (got == zo.g1_to_affine(zo.msm(pts, sc))).all()


def test_msm_edge_scalars(gpu, zo):
    """zero / one / r-1 scalars, signed-digit boundaries, repeated and opposite
    points (doubling and cancellation inside one bucket)"""
    c = 10
    half = 1 << (c - 1)
    boundary = sum(half << (c * w) for w in range(25)) % R
    ints = [0, 1, R - 1, 2, boundary, (boundary + 1) % R, R - 2, half, half + 1, half - 1] * 30
    n = len(ints)
    sc = zo.to_mont(ints)
    base = zo.gen_bases_random(8, 77)
    pts = np.concatenate([base[np.arange(n) % 3]])  # repeated points
    # opposite points: P and -P with equal scalars cancel
    neg = base[3:4].copy()
    from_mont_y = zo.g1a_to_ints(neg)[1]
    negp = np.concatenate([neg[:, :4], zo.fp_to_mont_int((zo.P_MODULUS - from_mont_y) % zo.P_MODULUS)], axis=1)
    pts = np.concatenate([pts, base[3:4], negp])
    sc = np.concatenate([sc, zo.to_mont([12345]), zo.to_mont([12345])])
    got = _affine_of(zo, _zk_msm(pts, sc, c))
    assert (got == zo.g1_to_affine(zo.msm_naive(pts, sc))).all()
    # all-zero scalars -> identity (0, 0)
    z = _affine_of(zo, _zk_msm(pts, np.zeros_like(sc)))
    assert (z == 0).all()


def _kat_inputs(zo, n, seed):
    """src/cuda/msm.rs:477-511: 256 candidate scalars c_j, bases c_{i mod 256} G"""
    cand = zo.rand_fr(256, seed)
    g = zo.g1_generator()
    cpts = zo.g1_to_affine(np.concatenate([zo.g1_mul(g, cand[j:j + 1]) for j in range(256)]))
    pts = np.ascontiguousarray(np.tile(cpts, (n // 256, 1)))
    return cand, g, pts


@pytest.mark.parametrize("log_n,batch", [(16, 1), (16, 3), (16, 8), (22, 1), (22, 2)])
def test_reference_known_answer(gpu, zo, log_n, batch):
    n = 1 << log_n
    cand, g, pts = _kat_inputs(zo, n, 4242)
    sets = [zo.rand_fr(n, 5000 + 17 * b + log_n) for b in range(batch)]
    got = _batch_msm(pts, sets, skip_zero=batch & 1)
    for b in range(batch):
        k = zo.fr_inner_cycled(sets[b], cand)
        exp = zo.g1_to_affine(zo.g1_mul(g, k))
        assert (got[b:b + 1] == exp).all(), b


@pytest.mark.parametrize("log_n,c", [(20, 6), (22, 10)])
def test_msm_large_sort_groups(gpu, zo, log_n, c):
    """windows whose sort groups exceed what one workgroup sorts whole (24576
    entries): 2^20 points at c = 6 give 32 groups of ~32k entries per window
    with a single low-bit key, 2^22 at c = 10 128 groups of ~32k; they are
    sorted tile by tile.  Checked by the known-answer construction."""
    n = 1 << log_n
    cand, g, pts = _kat_inputs(zo, n, 4545 + c)
    sc = zo.rand_fr(n, 7100 + c)
    got = zo.g1_to_affine(_zk_msm(pts, sc, c))
    assert (got == zo.g1_to_affine(zo.g1_mul(g, zo.fr_inner_cycled(sc, cand)))).all()


@pytest.mark.parametrize("n,batch", [(1000, 1), (1 << 14, 3)])
def test_device_and_host_tails_agree(gpu, zo, monkeypatch, n, batch):
    """the window combination on host cores (default) and on the GPU
    (ZK_MSM_TAIL=device) give the same canonical point"""
    pts = zo.gen_bases_ap(n, 61)
    sets = [zo.rand_fr(n, 62 + b) for b in range(batch)]
    exp = [zo.g1_to_affine(zo.msm(pts, s)) for s in sets]
    host = _batch_msm(pts, sets)
    monkeypatch.setenv("ZK_MSM_TAIL", "device")
    dev = _batch_msm(pts, sets)
    one = _affine_of(zo, _zk_msm(pts, sets[0]))
    monkeypatch.delenv("ZK_MSM_TAIL")
    for b in range(batch):
        assert (host[b:b + 1] == exp[b]).all()
        assert (dev[b:b + 1] == exp[b]).all()
    assert (one == exp[0]).all()


@pytest.mark.parametrize("c", [0, 5, 9, 13])
def test_msm_batch_ragged_matches_oracle(gpu, zo, c):
    """a batch of MSMs of unequal sizes (1 point, a partial sort tile, several
    tiles), one with all-zero scalars and two sharing their points, through
    one pass of the pipeline; each result equals the oracle's MSM"""
    sizes = [1000, 1, 4097, 1 << 14, 300, 1 << 14]
    pts = [zo.gen_bases_random(n, 70 + k) for k, n in enumerate(sizes)]
    pts[5] = pts[3]
    sc = [zo.rand_fr(n, 80 + k) for k, n in enumerate(sizes)]
    sc[4] = np.zeros_like(sc[4])
    got = zo.g1_to_affine(_zk_msm_batch(pts, sc, c))
    for m in range(len(sizes)):
        exp = zo.g1_to_affine(zo.msm(pts[m], sc[m]))
        assert (got[m:m + 1] == exp).all(), (m, sizes[m])
    assert (got[4] == 0).all()


def test_msm_batch_known_answer(gpu, zo):
    """the reference's known-answer construction for 6 MSMs of 2^18 points
    sharing the bases, in one batched pass"""
    n, batch = 1 << 18, 6
    cand, g, pts = _kat_inputs(zo, n, 4343)
    sets = [zo.rand_fr(n, 6100 + b) for b in range(batch)]
    got = zo.g1_to_affine(_zk_msm_batch([pts] * batch, sets))
    for b in range(batch):
        exp = zo.g1_to_affine(zo.g1_mul(g, zo.fr_inner_cycled(sets[b], cand)))
        assert (got[b:b + 1] == exp).all(), b


def test_msm_batch_equals_single_calls(gpu, zo, monkeypatch):
    """batched results equal zk_msm one by one, with the host and the device tail"""
    sizes = [1 << 12, 3000, 1 << 13]
    pts = [zo.gen_bases_ap(n, 90 + k) for k, n in enumerate(sizes)]
    sc = [zo.rand_fr(n, 95 + k) for k, n in enumerate(sizes)]
    one = [zo.g1_to_affine(_zk_msm(p, s, 10)) for p, s in zip(pts, sc)]
    host = zo.g1_to_affine(_zk_msm_batch(pts, sc, 10))
    monkeypatch.setenv("ZK_MSM_TAIL", "device")
    dev = zo.g1_to_affine(_zk_msm_batch(pts, sc, 10))
    monkeypatch.delenv("ZK_MSM_TAIL")
    for m in range(len(sizes)):
        assert (host[m:m + 1] == one[m]).all()
        assert (dev[m:m + 1] == one[m]).all()


@pytest.mark.parametrize("log_n,values", [(16, [1]), (16, [0, 1, 2, 3]), (18, [1, 2]), (14, [R - 1, 1])])
def test_msm_heavy_buckets(gpu, zo, log_n, values):
    """scalars from a tiny set put almost every entry into a few buckets that
    span thousands of accumulation chunks (the multiplicities of a lookup look
    like this); their partial sums are merged by one workgroup per bucket"""
    n = 1 << log_n
    rng = np.random.default_rng(log_n + len(values))
    pts = zo.gen_bases_random(n, 70 + log_n)
    sc = zo.to_mont([int(values[i]) for i in rng.integers(0, len(values), n)])
    exp = zo.g1_to_affine(zo.msm(pts, sc))
    assert (_affine_of(zo, _zk_msm(pts, sc)) == exp).all()
    got = zo.g1_to_affine(_zk_msm_batch([pts, pts], [sc, sc]))
    assert (got[0:1] == exp).all() and (got[1:2] == exp).all()


def _zk_msm_dev(dp, ds, n, c=0):
    """zk_msm over device arrays (torch tensors / pointer offsets in bytes)"""
    import torch
    from zkprover import device as dv
    from zkprover._lib import call, lib
    need = lib().zk_msm_scratch_bytes
    need.restype = ctypes.c_size_t
    need.argtypes = [ctypes.c_int, ctypes.c_int]
    nbytes = need(n, c)
    scratch = torch.empty(nbytes, dtype=torch.uint8, device=dv.device())
    out = torch.zeros(OUT_BYTES, dtype=torch.uint8, device=dv.device())
    lib().zk_msm.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                                      ctypes.c_size_t, ctypes.c_void_p]
    call("zk_msm", out.data_ptr(), dp, ds, n, c, scratch.data_ptr(), nbytes, dv.stream_handle())
    return out[:96].cpu().numpy().view(np.uint64).reshape(1, 12).copy()


def test_msm_two_entry_layouts_agree(gpu, zo):
    """above 2^23 scalars the sort passes (index, low bits) as two words,
    below as one packed word: an MSM of 2^23 + 3 points (two-word layout)
    equals the sum of the MSMs of its two halves (one-word layout)"""
    import torch
    from zkprover import device as dv
    from zkprover._lib import call, lib
    n = (1 << 23) + 3
    lib().zk_gen_points.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_void_p]
    lib().zk_rand_fr.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_void_p]
    pts = torch.empty(n * 64, dtype=torch.uint8, device=dv.device())
    sc = torch.empty(n * 32, dtype=torch.uint8, device=dv.device())
    call("zk_gen_points", pts.data_ptr(), n, 91, dv.stream_handle())
    call("zk_rand_fr", sc.data_ptr(), n, 92, dv.stream_handle())
    h = 1 << 22
    full = _zk_msm_dev(pts.data_ptr(), sc.data_ptr(), n)
    a = _zk_msm_dev(pts.data_ptr(), sc.data_ptr(), h)
    b = _zk_msm_dev(pts.data_ptr() + h * 64, sc.data_ptr() + h * 32, n - h)
    assert (zo.g1_to_affine(full) == zo.g1_to_affine(zo.g1_add(a, b))).all()


def test_collect_of_unfinished_slice_is_an_error(gpu, zo, monkeypatch):
    """batch_msm_collect over a slice msm() never finished (its header
    cleared) reports hipErrorInvalidValue instead of a commitment; the device
    tail marks such a slice with 96 bytes of 0xff (no point)"""
    import torch
    from zkprover import device as dv
    from zkprover._lib import HipError, call
    n = 1000
    pts = zo.gen_bases_random(n, 17)
    sc = zo.rand_fr(n, 18)
    good = _zk_msm(pts, sc)  # a finished slice for comparison
    assert (zo.g1_to_affine(good) == zo.g1_to_affine(zo.msm(pts, sc))).all()
    slice_buf = torch.zeros(OUT_BYTES, dtype=torch.uint8, device=dv.device())
    arr = dv.to_dev(np.array([[slice_buf.data_ptr()]], dtype=np.uint64))
    with pytest.raises(HipError):
        call("batch_msm_collect", None, None, None, None, arr.data_ptr(), 1, 1, 1, 1, dv.stream_handle())
    monkeypatch.setenv("ZK_MSM_TAIL", "device")
    call("batch_msm_collect", None, None, None, None, arr.data_ptr(), 1, 1, 1, 1, dv.stream_handle())
    torch.cuda.synchronize()
    monkeypatch.delenv("ZK_MSM_TAIL")
    assert (slice_buf[:96].cpu().numpy() == 0xff).all()


def test_msm_phase_profile_ring(gpu, zo):
    """70 profiled MSMs (more than the profiler's 64 event sets, so the ring
    is drained once while calls go on): reading returns all 70 calls with
    positive phase times, and profiling leaves the results unchanged"""
    from zkprover._lib import lib
    L = lib()
    L.zk_msm_profile_enable.argtypes = [ctypes.c_int]
    L.zk_msm_profile_enable.restype = None
    L.zk_msm_profile_read.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.zk_msm_profile_read.restype = ctypes.c_int
    n = 4096
    pts = zo.gen_bases_random(n, 61)
    sc = zo.rand_fr(n, 62)
    plain = zo.g1_to_affine(_zk_msm(pts, sc))
    times = np.zeros(8, dtype=np.float64)
    L.zk_msm_profile_read(times.ctypes.data, 8, 1)
    L.zk_msm_profile_enable(1)
    try:
        got = [zo.g1_to_affine(_zk_msm(pts, sc)) for _ in range(70)]
    finally:
        L.zk_msm_profile_enable(0)
    calls = L.zk_msm_profile_read(times.ctypes.data, 8, 1)
    assert calls == 70
    assert (times[:7] > 0).all()
    assert all((g == plain).all() for g in got)
    assert L.zk_msm_profile_read(times.ctypes.data, 8, 1) == 0
