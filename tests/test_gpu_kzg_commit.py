"""GPU parity: plain KZG commitment of one polynomial (2^16), the small
end-to-end case of the path: Lagrange values -> iNTT (w^-1 tables, divisor
2^-16, src/cuda/ntt.rs:61-93) -> MSM against the SRS (msm + batch_msm_collect
with the reference driver's buffers, src/cuda/msm.rs:239-455).  The point must
equal the CPU oracle's iNTT + MSM bit for bit (affine bytes).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("log_n", [12, 16])
def test_kzg_commit_matches_oracle(gpu, zo, log_n):
    from zkprover import device as dv
    from zkprover._lib import call, check, lib
    n = 1 << log_n
    values = zo.rand_fr(n, 1)            # p(X) in Lagrange form
    srs = zo.gen_bases_random(n, 2)      # g_i = k_i G, k_i seeded uniform
    w = zo.root_of_unity(log_n)
    w_inv = zo.fr_inv(w)
    n_inv = zo.fr_inv(zo.fr_from_u64(n))

    # device iNTT through the drop-in entries
    init = np.zeros((n, 4), dtype=np.uint64)
    init[0] = zo.fr_from_u64(1)[0]
    init[1] = w_inv[0]
    om = dv.to_dev(init)
    call("expand_omega_buffer", om.data_ptr(), log_n, 1)
    buf, tmp, dninv = dv.to_dev(values), dv.empty(n), dv.to_dev(n_inv)
    swap = ctypes.c_bool(False)
    call("ntt", buf.data_ptr(), tmp.data_ptr(), None, om.data_ptr(), dninv.data_ptr(), log_n, 8,
         ctypes.addressof(swap), dv.stream_handle())
    coeffs_dev = tmp if swap.value else buf

    # device MSM of the coefficients against the SRS (the reference's host driver path)
    dp = dv.to_dev(srs.reshape(-1, 4))
    arr = (ctypes.c_void_p * 1)(coeffs_dev.data_ptr())
    out = np.zeros((1, 8), dtype=np.uint64)
    f = lib().zkp_batch_msm
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    f.restype = ctypes.c_int
    check(f(out.ctypes.data, dp.data_ptr(), ctypes.cast(arr, ctypes.c_void_p), 1, n, 0), "zkp_batch_msm")

    coeffs = zo.ntt(values, log_n, w_inv, n_inv)
    assert (dv.from_dev(coeffs_dev) == coeffs).all()
    exp = zo.g1_to_affine(zo.msm(srs, coeffs))
    assert (out == exp).all()
    assert zo.g1a_is_on_curve(out)
