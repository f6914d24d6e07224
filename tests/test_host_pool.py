"""CPU: the pinned pool's bookkeeping (zkwasm-prover_amd/csrc/host_pool.h),
compiled with g++ over a backend of fake, never-dereferenced address ranges.

Checked after every step of seeded random alloc/free sequences, against the
live set the test keeps itself: live ranges are disjoint and sit at a slab
base or its one chunk boundary; a chunk request goes into a slab with one
live chunk whenever one exists, and the pool grows only when no slab has
room; a slab request takes an empty slab and grows only when none is; sizes
0 and above a slab, unknown pointers, double frees and interior pointers are
refused; a refusing backend leaves the state unchanged; the pool releases
every slab when destroyed.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BASE = 1 << 40


@pytest.fixture(scope="module")
def hp(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("hp") / "libhp.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I",
                    os.path.join(ROOT, "zkwasm-prover_amd", "csrc"),
                    os.path.join(HERE, "host_pool", "check.cpp"), "-o", out], check=True)
    L = ctypes.CDLL(out)
    L.hp_new.restype = ctypes.c_void_p
    L.hp_new.argtypes = [ctypes.c_long]
    L.hp_delete.argtypes = [ctypes.c_void_p]
    L.hp_reserve.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    L.hp_alloc.restype = ctypes.c_void_p
    L.hp_alloc.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    L.hp_free.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.hp_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    L.hp_backend_allocs.restype = ctypes.c_long
    L.hp_backend_releases.restype = ctypes.c_long
    L.hp_chunk.restype = ctypes.c_size_t
    L.hp_slab.restype = ctypes.c_size_t
    return L


class Model:
    """the live set as the test sees it; slab i of the fake backend sits at
    BASE + i * 2 * slab (each range is followed by a gap of one slab)"""

    def __init__(self, L, limit=-1):
        self.L, self.chunk, self.slab = L, L.hp_chunk(), L.hp_slab()
        self.h = L.hp_new(limit)
        self.live = {}  # address -> "c" | "s"

    def where(self, p):
        off = p - BASE
        return off // (2 * self.slab), off % (2 * self.slab)

    def slabs(self):
        s, u = ctypes.c_size_t(), ctypes.c_size_t()
        self.L.hp_stats(self.h, ctypes.byref(s), ctypes.byref(u))
        return s.value, u.value

    def occupancy(self):
        nslab = self.slabs()[0]
        occ = [[] for _ in range(nslab)]
        for p, kind in self.live.items():
            i, off = self.where(p)
            occ[i].append((off, kind))
        return occ

    def check(self):
        nslab, used = self.slabs()
        assert nslab == self.L.hp_backend_allocs()
        assert used == sum(1 if k == "c" else 2 for k in self.live.values())
        for i, cells in enumerate(self.occupancy()):
            offs = [o for o, _ in cells]
            assert len(offs) == len(set(offs)), "two allocations at one address"
            if any(k == "s" for _, k in cells):
                assert cells == [(0, "s")], "a whole-slab allocation shares its slab"
            else:
                assert all(o in (0, self.chunk) for o in offs) and len(cells) <= 2

    def alloc(self, nbytes):
        before = self.occupancy()
        p = self.L.hp_alloc(self.h, nbytes)
        if nbytes == 0 or nbytes > self.slab:
            assert not p
            return None
        grew = self.slabs()[0] > len(before)
        if not p:
            assert not grew
            return None
        i, off = self.where(p)
        if nbytes <= self.chunk:
            half = [j for j, c in enumerate(before) if len(c) == 1 and c[0][1] == "c"]
            empty = [j for j, c in enumerate(before) if not c]
            if half:
                assert i in half, "a chunk went past a half-used slab"
            elif empty:
                assert i in empty and not grew
            else:
                assert grew and i == len(before)
            self.live[p] = "c"
        else:
            empty = [j for j, c in enumerate(before) if not c]
            assert off == 0
            if empty:
                assert i in empty and not grew
            else:
                assert grew and i == len(before)
            self.live[p] = "s"
        self.check()
        return p

    def free(self, p):
        assert self.L.hp_free(self.h, p) == 1
        del self.live[p]
        self.check()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_sequences(hp, seed):
    m = Model(hp)
    rng = np.random.default_rng(seed)
    sizes = [1, 4096, m.chunk - 1, m.chunk, m.chunk + 1, m.slab - 1, m.slab, m.slab + 1, 0]
    for _ in range(1500):
        if m.live and rng.random() < 0.45:
            m.free(list(m.live)[rng.integers(len(m.live))])
        else:
            m.alloc(sizes[rng.integers(len(sizes))])
    for p in list(m.live):
        m.free(p)
    assert m.slabs()[1] == 0
    n = hp.hp_backend_allocs()
    hp.hp_delete(m.h)
    assert hp.hp_backend_releases() == n


def test_refusals(hp):
    m = Model(hp)
    a = m.alloc(m.slab)
    assert hp.hp_free(m.h, a + m.chunk) == 0, "interior pointer"
    assert hp.hp_free(m.h, BASE - 4096) == 0, "unknown pointer"
    m.free(a)
    assert hp.hp_free(m.h, a) == 0, "double free"
    c = m.alloc(100)
    d = m.alloc(100)
    assert m.where(c)[0] == m.where(d)[0] and {m.where(c)[1], m.where(d)[1]} == {0, m.chunk}
    m.free(c)
    assert hp.hp_free(m.h, c) == 0
    m.check()
    hp.hp_delete(m.h)


def test_fragmentation_prefers_half_used_slab(hp):
    m = Model(hp)
    a, b = m.alloc(10), m.alloc(10)  # slab 0 full
    m.free(a)                          # slab 0 half used
    s = m.alloc(m.slab)                # needs an empty slab: grows to slab 1
    assert m.where(s)[0] == 1
    m.free(s)                          # slab 1 empty, slab 0 half used
    c = m.alloc(10)
    assert m.where(c)[0] == 0, "the chunk must fill the half-used slab"
    assert m.slabs()[0] == 2
    hp.hp_delete(m.h)


def test_backend_refusal_and_reserve(hp):
    m = Model(hp, limit=2)
    assert hp.hp_reserve(m.h, 2) == 1
    assert m.slabs() == (2, 0)
    s1, s2 = m.alloc(m.slab), m.alloc(m.slab)
    assert s1 and s2
    assert m.alloc(10) is None        # backend refuses a third slab
    assert m.slabs() == (2, 4)
    assert hp.hp_reserve(m.h, 3) == 0
    m.free(s1)
    assert m.alloc(10) is not None
    hp.hp_delete(m.h)
    assert hp.hp_backend_releases() == 2
