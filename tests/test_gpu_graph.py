"""GPU: a round captured as a HIP graph (input copies, a batched NTT and a
batched MSM with the device tail) replays to the same bytes as the eager
sequence, every time, with unrelated eager work and synchronisations between
the replays; and the eager sequence itself matches the oracle."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OUT = 64 * 1024


class _NttItem(ctypes.Structure):
    _fields_ = [("buf", ctypes.c_void_p), ("tmp", ctypes.c_void_p), ("omegas", ctypes.c_void_p),
                ("n_inv", ctypes.c_void_p), ("log_n", ctypes.c_int), ("reserved", ctypes.c_int)]


class _MsmItem(ctypes.Structure):
    _fields_ = [("points", ctypes.c_void_p), ("scalars", ctypes.c_void_p), ("n", ctypes.c_int),
                ("reserved", ctypes.c_int)]


def test_graph_replay_matches_eager(gpu, zo):
    import torch
    from zkprover import device as dv
    from zkprover._lib import call, lib
    from zkprover.autotune import omega_table
    L = lib()
    L.zk_msm_batch_scratch_bytes.restype = ctypes.c_size_t
    L.zk_msm_batch_scratch_bytes.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream

    logs = [10, 11, 12, 13]
    xs = [zo.rand_fr(1 << k, 800 + k) for k in logs]
    orig = [dv.to_dev(x) for x in xs]
    bufs = [dv.empty(1 << k) for k in logs]
    tmps = [dv.empty(1 << k) for k in logs]
    tables = [omega_table(k) for k in logs]
    nitems = (_NttItem * len(logs))()
    for m, k in enumerate(logs):
        nitems[m] = _NttItem(bufs[m].data_ptr(), tmps[m].data_ptr(), tables[m].data_ptr(), None, k, 0)
    swapped = (ctypes.c_ubyte * len(logs))()

    sizes = [1000, 1 << 12, 5000]
    pts = [zo.gen_bases_random(n, 810 + i) for i, n in enumerate(sizes)]
    scs = [zo.rand_fr(n, 820 + i) for i, n in enumerate(sizes)]
    dpts = [dv.to_dev(p.reshape(-1, 4)) for p in pts]
    dscs = [dv.to_dev(s) for s in scs]
    mitems = (_MsmItem * len(sizes))()
    for m, n in enumerate(sizes):
        mitems[m] = _MsmItem(dpts[m].data_ptr(), dscs[m].data_ptr(), n, 0)
    nbytes = L.zk_msm_batch_scratch_bytes(ctypes.addressof(mitems), len(sizes), 0)
    scratch = torch.empty(nbytes, dtype=torch.uint8, device=dv.device())
    out = torch.zeros(len(sizes) * OUT, dtype=torch.uint8, device=dv.device())

    def round_():
        for m, k in enumerate(logs):
            call("zk_copy", bufs[m].data_ptr(), orig[m].data_ptr(), (1 << k) * 32, sh)
        call("zk_ntt_batch", ctypes.addressof(nitems), len(logs), ctypes.addressof(swapped), sh)
        call("zk_msm_batch_async", out.data_ptr(), ctypes.addressof(mitems), len(sizes), 0, scratch.data_ptr(),
             nbytes, sh)

    def results():
        torch.cuda.synchronize()
        ntts = [dv.from_dev(tmps[m] if swapped[m] else bufs[m]) for m in range(len(logs))]
        msm = out.view(len(sizes), OUT)[:, :96].cpu().numpy().copy().view(np.uint64).reshape(len(sizes), 12)
        return ntts, msm

    round_()
    eager_ntt, eager_msm = results()
    for m, k in enumerate(logs):
        assert (eager_ntt[m] == zo.ntt(xs[m], k, zo.root_of_unity(k))).all()
    for m in range(len(sizes)):
        assert (zo.g1_to_affine(eager_msm[m:m + 1]) == zo.g1_to_affine(zo.msm(pts[m], scs[m]))).all()

    exec_ = ctypes.c_void_p()
    call("zk_graph_begin", sh)
    round_()
    call("zk_graph_end", sh, ctypes.byref(exec_))
    try:
        other = dv.to_dev(zo.rand_fr(4096, 830))
        for rep in range(3):
            # clear the outputs eagerly, run unrelated eager work, then replay
            for b in bufs + tmps:
                b.zero_()
            out.zero_()
            call("zk_fr_op", other.data_ptr(), other.data_ptr(), other.data_ptr(), 1, 4096, dv.stream_handle())
            torch.cuda.synchronize()
            call("zk_graph_launch", exec_, sh)
            ntt_r, msm_r = results()
            for m in range(len(logs)):
                assert (ntt_r[m] == eager_ntt[m]).all(), (rep, m)
            assert (msm_r == eager_msm).all(), rep
    finally:
        call("zk_graph_destroy", exec_)
