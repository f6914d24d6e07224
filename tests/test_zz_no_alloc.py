"""GPU: zk_msm_noalloc (scratch in the library's device array) against zk_msm
(caller scratch), bit-exact, on two calls in a row, the second smaller, so
the second call runs over the first call's leftovers in the shared array.
Also: sizes and windows whose scratch exceeds the array are refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OUT_BYTES = 64 * 1024
MAX_N = 1 << 17


def _lib():
    from zkprover._lib import lib
    L = lib()
    L.zk_msm_noalloc.restype = ctypes.c_int
    L.zk_msm_noalloc.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    return L


def _noalloc(points_dev, scalars_dev, n, c=0):
    import torch
    from zkprover import device as dv
    out = torch.zeros(OUT_BYTES, dtype=torch.uint8, device=dv.device())
    err = _lib().zk_msm_noalloc(out.data_ptr(), points_dev.data_ptr(), scalars_dev.data_ptr(), n, c,
                                dv.stream_handle())
    assert err == 0, f"zk_msm_noalloc returned {err}"
    torch.cuda.synchronize()
    return out[:96].cpu().numpy().view(np.uint64).reshape(1, 12).copy()


def _with_scratch(points_dev, scalars_dev, n, c=0):
    import torch
    from zkprover import device as dv
    from zkprover._lib import call, lib
    L = lib()
    L.zk_msm_scratch_bytes.restype = ctypes.c_size_t
    L.zk_msm_scratch_bytes.argtypes = [ctypes.c_int, ctypes.c_int]
    L.zk_msm.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                                 ctypes.c_void_p]
    nbytes = L.zk_msm_scratch_bytes(n, c)
    scratch = torch.empty(nbytes, dtype=torch.uint8, device=dv.device())
    out = torch.zeros(OUT_BYTES, dtype=torch.uint8, device=dv.device())
    call("zk_msm", out.data_ptr(), points_dev.data_ptr(), scalars_dev.data_ptr(), n, c, scratch.data_ptr(), nbytes,
         dv.stream_handle())
    return out[:96].cpu().numpy().view(np.uint64).reshape(1, 12).copy()


@pytest.mark.parametrize("sizes", [(MAX_N, 1000), (4096, 3)])
def test_noalloc_equals_zk_msm_twice(gpu, zo, sizes):
    from zkprover import device as dv
    pts = zo.gen_bases_random(sizes[0], 11)
    for call_no, n in enumerate(sizes):
        sc = zo.rand_fr(n, 100 + call_no)
        dp, ds = dv.to_dev(pts[:n].reshape(-1, 4)), dv.to_dev(sc)
        print(f"zk_msm_noalloc call {call_no} n={n}", flush=True)
        got = _noalloc(dp, ds, n)
        want = _with_scratch(dp, ds, n)
        assert (got == want).all(), f"call {call_no} (n={n}): noalloc differs from zk_msm"
        if n <= 4096:
            assert (zo.g1_to_affine(got) == zo.g1_to_affine(zo.msm(pts[:n], sc))).all()


def test_noalloc_rejects_sizes_beyond_the_array(gpu):
    import torch
    from zkprover import device as dv
    out = torch.zeros(OUT_BYTES, dtype=torch.uint8, device=dv.device())
    dummy = torch.zeros(64, dtype=torch.uint8, device=dv.device())
    L = _lib()
    # the inputs are never read: the size check comes first
    assert L.zk_msm_noalloc(out.data_ptr(), dummy.data_ptr(), dummy.data_ptr(), 1 << 20, 0,
                            dv.stream_handle()) != 0
    assert L.zk_msm_noalloc(out.data_ptr(), dummy.data_ptr(), dummy.data_ptr(), MAX_N, 3,
                            dv.stream_handle()) != 0
    assert L.zk_msm_noalloc(out.data_ptr(), dummy.data_ptr(), dummy.data_ptr(), 0, 0, dv.stream_handle()) != 0
