"""Host residency: the pinned host pool and the device-to-host copy queue
(zk_pinned_* / zk_copyq_* of the kernel library; the reference's
src/hugetlb.rs and src/async_copy_queue.rs).

`PinnedBuffer` is one pool allocation seen as a numpy array; `CopyQueue`
orders device-to-host copies after a producer stream's work without host
waits, at most `max_inflight` outstanding.
"""
import ctypes

import numpy as np

from ._lib import HipError, check, lib

SNAPSHOT = 1
RELEASE = 2
CHUNK_BYTES = 1 << 27
SLAB_BYTES = 1 << 28


def _decl():
    L = lib()
    P, SZ = ctypes.c_void_p, ctypes.c_size_t
    for name, args in (("zk_pinned_reserve", [SZ]), ("zk_pinned_alloc", [SZ, P]), ("zk_pinned_free", [P]),
                       ("zk_pinned_stats", [P, P, P, P]), ("zk_copyq_create", [ctypes.c_int, P]),
                       ("zk_copyq_push", [P, P, P, SZ, P, ctypes.c_int]), ("zk_copyq_sync_all", [P]),
                       ("zk_copyq_stats", [P, P, P, P]), ("zk_copyq_destroy", [P]),
                       ("zk_prepare_advice_buffer", [ctypes.c_int, ctypes.c_int, P]),
                       ("zk_release_advice_buffer", [ctypes.c_int, P])):
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    return L


def reserve(slabs):
    check(_decl().zk_pinned_reserve(slabs), "zk_pinned_reserve")


def pool_stats():
    """(slabs, chunks in use)"""
    s, u = ctypes.c_size_t(), ctypes.c_size_t()
    check(_decl().zk_pinned_stats(ctypes.byref(s), ctypes.byref(u), None, None), "zk_pinned_stats")
    return s.value, u.value


class PinnedBuffer:
    def __init__(self, nbytes):
        self.L = _decl()
        self.nbytes = nbytes
        p = ctypes.c_void_p()
        check(self.L.zk_pinned_alloc(nbytes, ctypes.byref(p)), "zk_pinned_alloc")
        self.ptr = p.value

    def array(self, dtype=np.uint64, shape=None):
        """a numpy view of the pinned bytes (valid until close)"""
        if not self.ptr:
            raise HipError("PinnedBuffer used after close")
        raw = np.ctypeslib.as_array((ctypes.c_uint8 * self.nbytes).from_address(self.ptr))
        v = raw.view(dtype)
        return v if shape is None else v.reshape(shape)

    def close(self):
        if self.ptr:
            check(self.L.zk_pinned_free(self.ptr), "zk_pinned_free")
            self.ptr = None


class AdviceBuffer:
    """prepare_advice_buffer (src/lib.rs:107-112): `count` zeroed pinned
    columns of `rows` Fr each; `ptrs` are the addresses the prover's host
    witness takes, `column(c)` a numpy (rows, 4) uint64 view of column c"""

    def __init__(self, rows, count):
        self.L = _decl()
        self.rows, self.count = rows, count
        self._cols = (ctypes.c_void_p * count)()
        check(self.L.zk_prepare_advice_buffer(rows, count, self._cols), "zk_prepare_advice_buffer")
        self.ptrs = [self._cols[c] for c in range(count)]

    def column(self, c):
        if not self.ptrs:
            raise HipError("AdviceBuffer used after close")
        nbytes = self.rows * 32
        raw = np.ctypeslib.as_array((ctypes.c_uint8 * nbytes).from_address(self.ptrs[c]))
        return raw.view(np.uint64).reshape(self.rows, 4)

    def close(self):
        if self.ptrs:
            check(self.L.zk_release_advice_buffer(self.count, self._cols), "zk_release_advice_buffer")
            self.ptrs = []


class CopyQueue:
    def __init__(self, max_inflight=28):
        self.L = _decl()
        h = ctypes.c_void_p()
        check(self.L.zk_copyq_create(max_inflight, ctypes.byref(h)), "zk_copyq_create")
        self.h = h.value

    def push(self, host_ptr, dev_ptr, nbytes, stream=0, flags=0):
        check(self.L.zk_copyq_push(self.h, host_ptr, dev_ptr, nbytes, stream, flags), "zk_copyq_push")

    def sync_all(self):
        check(self.L.zk_copyq_sync_all(self.h), "zk_copyq_sync_all")

    def stats(self):
        """(pushed, pushes that waited for a full queue, longest such wait in ms)"""
        p, w, ms = ctypes.c_ulonglong(), ctypes.c_ulonglong(), ctypes.c_double()
        check(self.L.zk_copyq_stats(self.h, ctypes.byref(p), ctypes.byref(w), ctypes.byref(ms)), "zk_copyq_stats")
        return p.value, w.value, ms.value

    def close(self):
        if self.h:
            check(self.L.zk_copyq_destroy(self.h), "zk_copyq_destroy")
            self.h = None


__all__ = ["PinnedBuffer", "CopyQueue", "reserve", "pool_stats", "SNAPSHOT", "RELEASE", "CHUNK_BYTES",
           "SLAB_BYTES"]
