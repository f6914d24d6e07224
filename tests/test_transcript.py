"""CPU: the Fiat-Shamir transcript (csrc/transcript.cpp).

Blake2b is checked against Python's hashlib (RFC 7693, with the
personalisation parameter) over lengths around the 128-byte block edges and
every digest size; the transcript against a model of halo2's Blake2bWrite
written with hashlib (point 0x01 || x || y, scalar 0x02 || s, challenge 0x00
then the digest of the running state read little-endian mod r).
"""
import ctypes
import hashlib

import numpy as np
import pytest

R = 21888242871839275222246405745257275088548364400416034343698204186575808495617
PERSONAL = b"Halo2-Transcript"


@pytest.fixture(scope="module")
def L():
    from zkprover._lib import lib
    l = lib()
    l.zk_blake2b.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    l.zk_blake2b.restype = ctypes.c_int
    l.zk_transcript_new.restype = ctypes.c_void_p
    l.zk_transcript_free.argtypes = [ctypes.c_void_p]
    l.zk_transcript_free.restype = None
    for f in ("zk_transcript_common_point", "zk_transcript_common_scalar", "zk_transcript_squeeze"):
        getattr(l, f).argtypes = [ctypes.c_void_p] * 3 if f != "zk_transcript_common_scalar" else [ctypes.c_void_p] * 2
        getattr(l, f).restype = ctypes.c_int
    return l


@pytest.mark.parametrize("n", [0, 1, 3, 63, 64, 127, 128, 129, 255, 256, 257, 1000, 4096])
@pytest.mark.parametrize("outlen", [64, 32, 1, 20])
@pytest.mark.parametrize("person", [None, PERSONAL])
def test_blake2b_matches_hashlib(L, n, outlen, person):
    data = np.random.default_rng(n * 7 + outlen).integers(0, 256, size=n, dtype=np.uint8).tobytes()
    out = ctypes.create_string_buffer(outlen)
    pers = ctypes.create_string_buffer(person, 16) if person else None
    assert L.zk_blake2b(out, outlen, data if n else None, n, pers) == 0
    kw = {"person": person} if person else {}
    assert out.raw == hashlib.blake2b(data, digest_size=outlen, **kw).digest()


def test_transcript_matches_model(L, zo):
    rng = np.random.default_rng(5)
    t = L.zk_transcript_new()
    h = hashlib.blake2b(digest_size=64, person=PERSONAL)
    try:
        for step in range(12):
            kind = step % 3
            if kind == 0:
                x, y = (int(v) % R for v in rng.integers(0, 2**62, size=2))
                xb, yb = x.to_bytes(32, "little"), y.to_bytes(32, "little")
                assert L.zk_transcript_common_point(t, xb, yb) == 0
                h.update(b"\x01" + xb + yb)
            elif kind == 1:
                s = (int(rng.integers(0, 2**62)) << 190) % R
                sb = s.to_bytes(32, "little")
                assert L.zk_transcript_common_scalar(t, sb) == 0
                h.update(b"\x02" + sb)
            else:
                mont = np.zeros((1, 4), dtype=np.uint64)
                dig = ctypes.create_string_buffer(64)
                assert L.zk_transcript_squeeze(t, mont.ctypes.data, dig) == 0
                h.update(b"\x00")
                d = h.copy().digest()
                assert dig.raw == d
                assert zo.from_mont(mont)[0] == int.from_bytes(d, "little") % R
    finally:
        L.zk_transcript_free(t)


def test_challenge_reduction_edge(L, zo):
    """digests near 2^512 exercise the unreduced halves of from_bytes_wide"""
    # drive the state until a digest has both halves >= r (most do: r < 2^254)
    t = L.zk_transcript_new()
    try:
        seen = 0
        for _ in range(8):
            mont = np.zeros((1, 4), dtype=np.uint64)
            dig = ctypes.create_string_buffer(64)
            L.zk_transcript_squeeze(t, mont.ctypes.data, dig)
            v = int.from_bytes(dig.raw, "little")
            lo, hi = v & (2**256 - 1), v >> 256
            seen += lo >= R and hi >= R
            assert zo.from_mont(mont)[0] == v % R
        assert seen > 0
    finally:
        L.zk_transcript_free(t)
