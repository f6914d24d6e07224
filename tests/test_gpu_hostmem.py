"""GPU: host residency (csrc/hostmem.cpp) through the C-ABI.

The pinned pool hands out page-locked memory that round-trips through the
device byte for byte and returns to the pool; the copy queue delivers every
pushed buffer's contents as they were when pushed: in snapshot mode even
though the producer's stream overwrites the buffer right after each push and
the queue is full most of the time, in the default mode for buffers left
alone, and in release mode for buffers the queue frees itself.  Malformed
pushes and pool requests are refused.  The prover writes its h pieces into
pinned memory with the same bytes as into pageable memory.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hm(gpu):
    from zkprover import hostmem
    return hostmem


@pytest.fixture(scope="module")
def hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return h


def _rand(n, seed):
    """the device generator's values for (n, seed), read back synchronously"""
    import torch
    from zkprover import device as dv
    from zkprover._lib import call
    t = dv.empty(n)
    call("zk_rand_fr", t.data_ptr(), n, seed, dv.stream_handle())
    torch.cuda.synchronize()
    return dv.from_dev(t)


def test_pinned_pool_roundtrip(hm, hip):
    import torch
    from zkprover import device as dv
    hm.reserve(1)
    slabs0, used0 = hm.pool_stats()
    assert slabs0 >= 1
    a = hm.PinnedBuffer(1 << 20)
    b = hm.PinnedBuffer(hm.CHUNK_BYTES)
    s = hm.PinnedBuffer(hm.SLAB_BYTES)
    assert hm.pool_stats()[1] == used0 + 4
    assert len({a.ptr, b.ptr, s.ptr}) == 3
    src = a.array(np.uint64)
    src[:] = np.arange(len(src), dtype=np.uint64) * np.uint64(0x9e3779b97f4a7c15)
    t = torch.empty(len(src), dtype=torch.int64, device=dv.device())
    assert hip.hipMemcpy(t.data_ptr(), a.ptr, a.nbytes, 1) == 0  # host to device
    torch.cuda.synchronize()
    dst = s.array(np.uint64)[:len(src)]
    assert hip.hipMemcpy(s.ptr, t.data_ptr(), a.nbytes, 2) == 0  # device to host
    assert (dst == src).all()
    L = hm._decl()
    p = ctypes.c_void_p()
    assert L.zk_pinned_alloc(hm.SLAB_BYTES + 1, ctypes.byref(p)) != 0
    assert L.zk_pinned_alloc(0, ctypes.byref(p)) != 0
    assert L.zk_pinned_free(a.ptr + 64) != 0
    for x in (a, b, s):
        x.close()
    assert hm.pool_stats()[1] == used0
    assert L.zk_pinned_free(b.ptr or 1) != 0


def test_copy_queue_snapshot_survives_overwrite(hm):
    import torch
    from zkprover import device as dv
    from zkprover._lib import call
    n, pushes, inflight = 1 << 16, 40, 8
    nbytes = n * 32
    host = hm.PinnedBuffer(pushes * nbytes)
    q = hm.CopyQueue(inflight)
    prod = torch.cuda.Stream()
    try:
        with torch.cuda.stream(prod):
            buf = dv.empty(n)
            h = prod.cuda_stream
            for i in range(pushes):
                call("zk_rand_fr", buf.data_ptr(), n, 1000 + i, h)
                q.push(host.ptr + i * nbytes, buf.data_ptr(), nbytes, h, hm.SNAPSHOT)
        q.sync_all()
        torch.cuda.synchronize()
        got = host.array(np.uint64, (pushes, n, 4))
        for i in range(pushes):
            assert (got[i] == _rand(n, 1000 + i)).all(), "push %d" % i
        pushed, waits, _ = q.stats()
        assert pushed == pushes and waits == pushes - inflight
    finally:
        q.close()
        host.close()


def test_copy_queue_default_and_release(hm, hip):
    import torch
    from zkprover import device as dv
    from zkprover._lib import call
    n, count = 1 << 15, 6
    nbytes = n * 32
    host = hm.PinnedBuffer(2 * count * nbytes)
    q = hm.CopyQueue(4)
    prod = torch.cuda.Stream()
    try:
        h = prod.cuda_stream
        with torch.cuda.stream(prod):
            keep = [dv.empty(n) for _ in range(count)]
        for i, t in enumerate(keep):  # default mode: buffers left alone until sync_all
            call("zk_rand_fr", t.data_ptr(), n, 2000 + i, h)
            q.push(host.ptr + i * nbytes, t.data_ptr(), nbytes, h, 0)
        for i in range(count):  # release mode: hipMalloc'd buffers the queue frees
            p = ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(p), nbytes) == 0
            call("zk_rand_fr", p.value, n, 3000 + i, h)
            q.push(host.ptr + (count + i) * nbytes, p.value, nbytes, h, hm.RELEASE)
        q.sync_all()
        torch.cuda.synchronize()
        got = host.array(np.uint64, (2 * count, n, 4))
        for i in range(count):
            assert (got[i] == _rand(n, 2000 + i)).all()
            assert (got[count + i] == _rand(n, 3000 + i)).all()
    finally:
        q.close()
        host.close()


def test_copy_queue_rejects_malformed_pushes(hm):
    import torch
    from zkprover import device as dv
    L = hm._decl()
    q = hm.CopyQueue(2)
    host = hm.PinnedBuffer(4096)
    t = dv.empty(128)
    try:
        for flags in (hm.SNAPSHOT | hm.RELEASE, 4):
            assert L.zk_copyq_push(q.h, host.ptr, t.data_ptr(), 4096, 0, flags) != 0
        assert L.zk_copyq_push(q.h, host.ptr, t.data_ptr(), 0, 0, 0) != 0
        assert L.zk_copyq_push(q.h, None, t.data_ptr(), 4096, 0, 0) != 0
        assert L.zk_copyq_create(0, ctypes.byref(ctypes.c_void_p())) != 0
        assert q.stats()[0] == 0
        torch.cuda.synchronize()
    finally:
        q.close()
        host.close()


def test_prover_pieces_into_pinned_memory(hm):
    from zkprover.prover import BATCHED, Circuit, Prover, ProvingKey
    circ = Circuit(k=10, advice=30, fixed=2, instance=1, lookups=3, table_log=7, perm_cols=6, shuffles=2,
                   shuffle_log=7, blinding=5)
    pk = ProvingKey(circ, seed=3)
    pr = Prover(pk)
    pin = hm.PinnedBuffer((1 << circ.k) * 32)
    try:
        com_a, pieces_a, _ = pr.prove(7, flags=BATCHED)
        out = pin.array(np.uint64, (1 << circ.k, 4))
        com_b, pieces_b, _ = pr.prove(7, flags=BATCHED, pieces_out=out)
        assert pieces_b is out
        assert (com_a == com_b).all() and (pieces_a == out).all()
    finally:
        pin.close()
        pr.close()
        pk.close()
