"""GPU parity: fixed-base MSM tables (zk_msm_table_build, zk_msm_item.table_c).

A table row T[w n + i] must be the affine point 2^(c w) P_i, checked against
the oracle's scalar multiplication.  An MSM over a table must equal the
oracle's MSM over the original points (bit-exact affine points) for windows
8..20, batches that mix plain and table items, identity and repeated points,
and the scalars at the signed-digit boundaries; the reference's known-answer
construction (src/cuda/msm.rs:457-537) pins the 2^22 case.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

R = 21888242871839275222246405745257275088548364400416034343698204186575808495617
OUT_BYTES = 64 * 1024


class _Item(ctypes.Structure):
    _fields_ = [("points", ctypes.c_void_p), ("scalars", ctypes.c_void_p), ("n", ctypes.c_int),
                ("table_c", ctypes.c_int)]


def _decl():
    from zkprover._lib import lib
    L = lib()
    L.zk_msm_table_bytes.restype = ctypes.c_size_t
    L.zk_msm_table_bytes.argtypes = [ctypes.c_int, ctypes.c_int]
    L.zk_msm_table_window.restype = ctypes.c_int
    L.zk_msm_table_window.argtypes = [ctypes.c_int]
    L.zk_msm_table_build.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    L.zk_msm_batch_scratch_bytes.restype = ctypes.c_size_t
    L.zk_msm_batch_scratch_bytes.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    for f in ("zk_msm_batch", "zk_msm_batch_async"):
        getattr(L, f).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                  ctypes.c_size_t, ctypes.c_void_p]
    return L


def _table(dp, n, c):
    """device table for the device points dp (torch tensor) at window c"""
    import torch
    from zkprover import device as dv
    from zkprover._lib import call
    L = _decl()
    nbytes = L.zk_msm_table_bytes(n, c)
    assert nbytes == -(-255 // c) * n * 64
    t = torch.empty(nbytes, dtype=torch.uint8, device=dv.device())
    call("zk_msm_table_build", t.data_ptr(), dp.data_ptr(), n, c, dv.stream_handle())
    return t


def _run_batch(specs, fn="zk_msm_batch", c=0):
    """specs: (points (n,8) host array or device tensor, scalars (n,4), table_c);
    table_c > 0 builds a table of the points first.  Returns (count, 12)."""
    import torch
    from zkprover import device as dv
    from zkprover._lib import call
    L = _decl()
    keep = []
    items = (_Item * len(specs))()
    for m, (p, s, tc) in enumerate(specs):
        dp = p if isinstance(p, torch.Tensor) else dv.to_dev(np.ascontiguousarray(p).reshape(-1, 4))
        ds = dv.to_dev(s)
        n = len(s)
        src = _table(dp, n, tc) if tc else dp
        keep += [dp, ds, src]
        items[m] = _Item(src.data_ptr(), ds.data_ptr(), n, tc)
    count = len(specs)
    nbytes = L.zk_msm_batch_scratch_bytes(ctypes.addressof(items), count, c)
    assert nbytes > 0
    scratch = torch.empty(nbytes, dtype=torch.uint8, device=dv.device())
    out = torch.zeros(count * OUT_BYTES, dtype=torch.uint8, device=dv.device())
    call(fn, out.data_ptr(), ctypes.addressof(items), count, c, scratch.data_ptr(), nbytes, dv.stream_handle())
    torch.cuda.synchronize()
    o = out.view(count, OUT_BYTES)[:, :96].cpu().numpy().copy()
    return o.view(np.uint64).reshape(count, 12)


def _with_identities(pts, where):
    pts = pts.copy()
    pts[where] = 0
    return pts


@pytest.mark.parametrize("c", [8, 13, 20])
def test_table_rows_are_window_multiples(gpu, zo, c):
    """T[w n + i] = 2^(c w) P_i as an affine point; identity points stay (0, 0)"""
    from zkprover import device as dv
    n = 257
    pts = _with_identities(zo.gen_bases_random(n, 300 + c), [0, 100, 256])
    dp = dv.to_dev(pts.reshape(-1, 4))
    tab = _table(dp, n, c).cpu().numpy().view(np.uint64).reshape(-1, 8)
    wd = -(-255 // c)
    assert tab.shape == (wd * n, 8)
    for w in (0, 1, wd // 2, wd - 1):
        k = zo.to_mont([pow(2, c * w, R)])
        rows = tab[w * n:(w + 1) * n]
        exp = zo.g1_to_affine(np.concatenate([zo.g1_mul(pts[i:i + 1], k) for i in range(n)]))
        assert (rows == exp).all(), w
        assert (rows[[0, 100, 256]] == 0).all()


@pytest.mark.parametrize("c", [8, 10, 12, 16, 19, 20])
@pytest.mark.parametrize("n", [1, 1000, 4097, 1 << 14])
def test_table_msm_matches_oracle(gpu, zo, c, n):
    pts = zo.gen_bases_random(n, 400 + n)
    sc = zo.rand_fr(n, 500 + n + c)
    got = zo.g1_to_affine(_run_batch([(pts, sc, c)]))
    assert (got == zo.g1_to_affine(zo.msm(pts, sc))).all()


@pytest.mark.parametrize("c", [11, 20])
def test_table_edge_scalars(gpu, zo, c):
    """zero / one / r-1 scalars, signed-digit boundaries of the table's window,
    repeated and opposite points (doubling and cancellation in one bucket, now
    across windows too), identity points; all-zero scalars give (0, 0)"""
    wd = -(-255 // c)
    half = 1 << (c - 1)
    boundary = sum(half << (c * w) for w in range(wd)) % R
    ints = [0, 1, R - 1, 2, boundary, (boundary + 1) % R, R - 2, half, half + 1, half - 1, 1 << c,
            (1 << c) - 1, pow(2, c * (wd - 1), R)] * 25
    n = len(ints)
    sc = zo.to_mont(ints)
    base = zo.gen_bases_random(8, 77)
    pts = base[np.arange(n) % 3].copy()
    neg = base[3:4].copy()
    y = zo.g1a_to_ints(neg)[1]
    negp = np.concatenate([neg[:, :4], zo.fp_to_mont_int((zo.P_MODULUS - y) % zo.P_MODULUS)], axis=1)
    pts = np.concatenate([pts, base[3:4], negp, np.zeros((2, 8), dtype=np.uint64)])
    sc = np.concatenate([sc, zo.to_mont([12345, 12345, 7, R - 1])])
    got = zo.g1_to_affine(_run_batch([(pts, sc, c)]))
    assert (got == zo.g1_to_affine(zo.msm_naive(pts, sc))).all()
    z = zo.g1_to_affine(_run_batch([(pts, np.zeros_like(sc), c)]))
    assert (z == 0).all()


@pytest.mark.parametrize("fn,tail", [("zk_msm_batch", "host"), ("zk_msm_batch", "device"),
                                     ("zk_msm_batch_async", "device")])
def test_table_batch_mixes_kinds(gpu, zo, monkeypatch, fn, tail):
    """plain items and table items of two windows in one call (one pass per
    kind, in one scratch); each result equals the oracle's MSM"""
    if tail == "device":
        monkeypatch.setenv("ZK_MSM_TAIL", "device")
    sizes = [1000, 1 << 14, 4097, 1, 1 << 14, 300]
    kinds = [0, 20, 13, 0, 20, 16]
    pts = [zo.gen_bases_random(n, 600 + k) for k, n in enumerate(sizes)]
    pts[4] = pts[1]
    sc = [zo.rand_fr(n, 700 + k) for k, n in enumerate(sizes)]
    sc[5] = np.zeros_like(sc[5])
    got = zo.g1_to_affine(_run_batch(list(zip(pts, sc, kinds)), fn=fn, c=9))
    for m in range(len(sizes)):
        exp = zo.g1_to_affine(zo.msm(pts[m], sc[m]))
        assert (got[m:m + 1] == exp).all(), (m, sizes[m], kinds[m])
    assert (got[5] == 0).all()


@pytest.mark.parametrize("log_n,batch", [(18, 3), (22, 1)])
def test_table_known_answer(gpu, zo, log_n, batch):
    """the reference's known-answer construction: bases c_{i mod 256} G,
    expected (sum_i s_i c_{i mod 256}) G, through tables at the window
    zk_msm_table_window picks"""
    import torch
    from zkprover import device as dv
    n = 1 << log_n
    cand = zo.rand_fr(256, 4444)
    g = zo.g1_generator()
    cpts = zo.g1_to_affine(np.concatenate([zo.g1_mul(g, cand[j:j + 1]) for j in range(256)]))
    dp = dv.to_dev(np.ascontiguousarray(np.tile(cpts, (n // 256, 1))).reshape(-1, 4))
    c = _decl().zk_msm_table_window(n)
    assert 8 <= c <= 20
    sets = [zo.rand_fr(n, 8000 + 17 * b + log_n) for b in range(batch)]
    got = zo.g1_to_affine(_run_batch([(dp, s, c) for s in sets]))
    for b in range(batch):
        exp = zo.g1_to_affine(zo.g1_mul(g, zo.fr_inner_cycled(sets[b], cand)))
        assert (got[b:b + 1] == exp).all(), b
    del dp
    torch.cuda.empty_cache()


def test_table_rejects_invalid(gpu, zo):
    """windows outside 8..20, n <= 0 and null pointers are errors; a batch with
    such an item has no scratch size and fails"""
    import torch
    from zkprover import device as dv
    from zkprover._lib import HipError, call
    L = _decl()
    assert L.zk_msm_table_bytes(100, 7) == 0 and L.zk_msm_table_bytes(100, 21) == 0
    assert L.zk_msm_table_bytes(0, 12) == 0
    buf = torch.zeros(13 * 64 * 4, dtype=torch.uint8, device=dv.device())
    with pytest.raises(HipError):
        call("zk_msm_table_build", buf.data_ptr(), buf.data_ptr(), 4, 21, dv.stream_handle())
    with pytest.raises(HipError):
        call("zk_msm_table_build", None, buf.data_ptr(), 4, 20, dv.stream_handle())
    items = (_Item * 1)(_Item(buf.data_ptr(), buf.data_ptr(), 4, 21))
    assert L.zk_msm_batch_scratch_bytes(ctypes.addressof(items), 1, 0) == 0
    out = torch.zeros(OUT_BYTES, dtype=torch.uint8, device=dv.device())
    with pytest.raises(HipError):
        call("zk_msm_batch", out.data_ptr(), ctypes.addressof(items), 1, 0, buf.data_ptr(), buf.numel(),
             dv.stream_handle())
