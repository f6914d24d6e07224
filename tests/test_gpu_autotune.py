"""GPU: the native autotuners time real candidates, record a choice inside
the tried range, and the tuned library still computes the same results
(every window / radix plan is exact, so tuning changes speed only)."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


def test_msm_autotune_then_exact(gpu, zo):
    from zkprover import autotune
    from zkprover._lib import lib
    L = lib()
    L.zk_msm_get_window.restype = ctypes.c_int
    try:
        res = autotune.tune_msm([12], reps=1)
        c, times = res[12]
        assert c in times and all(t > 0 for t in times.values())
        assert L.zk_msm_get_window(12) == c
        # an MSM of that class now runs at the tuned window and stays exact
        from test_gpu_msm import _zk_msm
        pts = zo.gen_bases_random(3000, 5)
        sc = zo.rand_fr(3000, 6)
        assert (zo.g1_to_affine(_zk_msm(pts, sc)) == zo.g1_to_affine(zo.msm(pts, sc))).all()
    finally:
        L.zk_msm_set_window(12, 0)


def test_ntt_autotune_then_exact(gpu, zo):
    from zkprover import autotune
    from zkprover._lib import lib
    from test_gpu_field_ntt import _gpu_ntt
    L = lib()
    L.zk_ntt_get_max_deg.restype = ctypes.c_int
    try:
        res = autotune.tune_ntt([14], reps=2)
        d, times = res[14]
        assert d in times and 4 <= d <= 8
        assert L.zk_ntt_get_max_deg(14) == d
        x = zo.rand_fr(1 << 14, 9)
        w = zo.root_of_unity(14)
        assert (_gpu_ntt(zo, x, 14, w) == zo.ntt(x, 14, w)).all()
    finally:
        L.zk_ntt_set_max_deg(14, 0)
