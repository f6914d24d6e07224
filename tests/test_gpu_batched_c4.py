"""GPU parity at the batched-primitives workload: 64 independent G1 MSMs with
sizes log-uniform in [2^12, 2^20] and 64 Fr NTTs with sizes 2^14..2^21,
uniform Fr data from fixed seeds, each through ONE batched call
(zk_msm_batch: more than 16 items, so the item table is consumed in several
groups per kernel launch; zk_ntt_batch: more than 48 transforms, so several
launches per pass round), every result bit-exact against the CPU oracle.

The reference's batch API takes one length for every MSM of a batch
(src/cuda/msm.rs:256-263), so its plain path issues uneven sizes one call at
a time; the batched results must equal those per-call results as well, which
the oracle's per-item MSM stands for here.  The oracle's MSM is pinned by the
naive sum and the reference's known-answer construction (test_oracle_golden).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OUT_BYTES = 64 * 1024


class _MsmItem(ctypes.Structure):
    _fields_ = [("points", ctypes.c_void_p), ("scalars", ctypes.c_void_p), ("n", ctypes.c_int),
                ("reserved", ctypes.c_int)]


class _NttItem(ctypes.Structure):
    _fields_ = [("buf", ctypes.c_void_p), ("tmp", ctypes.c_void_p), ("omegas", ctypes.c_void_p),
                ("n_inv", ctypes.c_void_p), ("log_n", ctypes.c_int), ("reserved", ctypes.c_int)]


def _c4_msm_sizes():
    rng = np.random.default_rng(8)
    return [int(round(2.0 ** x)) for x in rng.uniform(12, 20, 64)]


def test_c4_64_uneven_msms_one_batch(gpu, zo):
    import torch
    from zkprover import device as dv
    from zkprover._lib import call, lib
    sizes = _c4_msm_sizes()
    assert len(sizes) == 64 and min(sizes) >= 1 << 12 and max(sizes) <= 1 << 20
    nb = 1 << 20
    bases = zo.gen_bases_ap(nb, 808)
    assert all(zo.g1a_is_on_curve(bases[i:i + 1]) for i in (0, 1, 2, 3, nb - 4, nb - 3, nb - 2, nb - 1))
    rng = np.random.default_rng(88)
    offs = [int(rng.integers(0, nb - k + 1)) for k in sizes]
    scal = [zo.rand_fr(k, 8000 + m) for m, k in enumerate(sizes)]
    d_bases = dv.to_dev(bases.reshape(-1, 4))
    d_scal = [dv.to_dev(s) for s in scal]
    items = (_MsmItem * 64)()
    for m, k in enumerate(sizes):
        items[m] = _MsmItem(d_bases.data_ptr() + offs[m] * 64, d_scal[m].data_ptr(), k, 0)
    need = lib().zk_msm_batch_scratch_bytes
    need.restype = ctypes.c_size_t
    need.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    nbytes = need(ctypes.addressof(items), 64, 0)
    assert nbytes > 0
    scratch = torch.empty(nbytes, dtype=torch.uint8, device=dv.device())
    out = torch.zeros(64 * OUT_BYTES, dtype=torch.uint8, device=dv.device())
    lib().zk_msm_batch.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                   ctypes.c_size_t, ctypes.c_void_p]
    call("zk_msm_batch", out.data_ptr(), ctypes.addressof(items), 64, 0, scratch.data_ptr(), nbytes,
         dv.stream_handle())
    got = zo.g1_to_affine(out.view(64, OUT_BYTES)[:, :96].cpu().numpy().copy().view(np.uint64).reshape(64, 12))
    bad = []
    for m, k in enumerate(sizes):
        exp = zo.g1_to_affine(zo.msm(bases[offs[m]:offs[m] + k], scal[m]))
        if not (got[m:m + 1] == exp).all():
            bad.append((m, k))
    assert not bad, bad


def test_c4_64_ntts_one_batch(gpu, zo):
    from zkprover import device as dv
    from zkprover._lib import call, lib
    rng = np.random.default_rng(9)
    logs = [int(x) for x in rng.integers(14, 22, 64)]
    assert min(logs) == 14 and max(logs) == 21
    tables, keep, bufs, exps = {}, [], [], []
    items = (_NttItem * 64)()
    for m, k in enumerate(logs):
        w = zo.root_of_unity(k)
        if k not in tables:
            init = np.zeros((1 << k, 4), dtype=np.uint64)
            init[0] = zo.fr_from_u64(1)[0]
            init[1] = w[0]
            t = dv.to_dev(init)
            call("expand_omega_buffer", t.data_ptr(), k, 1)
            tables[k] = t
        x = zo.rand_fr(1 << k, 9000 + m)
        inv = m % 4 == 3  # every fourth transform an iNTT (w^-1 table, n^-1 scaling)
        if inv:
            wi = zo.fr_inv(w)
            key = -k
            if key not in tables:
                init = np.zeros((1 << k, 4), dtype=np.uint64)
                init[0] = zo.fr_from_u64(1)[0]
                init[1] = wi[0]
                t = dv.to_dev(init)
                call("expand_omega_buffer", t.data_ptr(), k, 1)
                tables[key] = t
            ninv = zo.fr_inv(zo.fr_from_u64(1 << k))
            exps.append(zo.ntt(x, k, wi, ninv))
            ni = dv.to_dev(ninv)
            keep.append(ni)
            om = tables[key]
        else:
            exps.append(zo.ntt(x, k, w))
            ni, om = None, tables[k]
        b, t = dv.to_dev(x), dv.empty(1 << k)
        bufs.append((b, t))
        items[m] = _NttItem(b.data_ptr(), t.data_ptr(), om.data_ptr(), ni.data_ptr() if ni is not None else None,
                            k, 0)
    swapped = (ctypes.c_ubyte * 64)()
    lib().zk_ntt_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    call("zk_ntt_batch", ctypes.addressof(items), 64, ctypes.addressof(swapped), dv.stream_handle())
    bad = [m for m, (b, t) in enumerate(bufs)
           if not (dv.from_dev(t if swapped[m] else b) == exps[m]).all()]
    assert not bad, [(m, logs[m]) for m in bad]
