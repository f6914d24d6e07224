"""GPU: the prover core's composition of every primitive (advice commitments
and iNTT, multiplicities, logup / permutation / shuffle grand products,
h(X) on the extended coset, vanishing) is checked by what the mathematics
fixes, independent of any restatement:
  * every grand product closes at row u (logup 0, permutation and shuffle 1);
  * h(X) = C(X) / Z_H(X) is a polynomial of degree < 3n, i.e. the extended
    iNTT's coefficients [3n, 4n) vanish -- which holds only if the gate, the
    permutation, logup and shuffle terms, the rotations, the coset and Z_H are
    all right;
  * a witness with one broken gate row fails that check (the check has teeth);
  * the batched-MSM and the per-polynomial paths give identical bytes, and a
    proof is a deterministic function of its seed;
  * the multi-open: with GWC each batch polynomial evaluates at its point to
    the v-combination of the evaluations (the division by X - z leaves that
    remainder); with Shplonk L(u) = 0, which holds only if every evaluation,
    the interpolations, h(X) and the divisions are right; one evaluation off
    by one fails either check.
The commitments themselves are MSMs whose parity is covered by the MSM tests.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _small(k=10):
    from zkprover.prover import Circuit
    return Circuit(k=k, advice=30, fixed=4, instance=1, lookups=3, table_log=7, perm_cols=6, shuffles=2,
                   shuffle_log=7, blinding=5)


@pytest.fixture(scope="module")
def core(gpu):
    from zkprover.prover import Prover, ProvingKey
    pk = ProvingKey(_small(), seed=3)
    pr = Prover(pk)
    yield pk, pr
    pr.close()
    pk.close()


def test_prover_core_constraints_hold(core, zo):
    pk, pr = core
    com, pieces, info = pr.prove(11)
    # instance, advice, m, permutation z (3 sets), logup z, shuffle z, the
    # random polynomial, quotient pieces, GWC quotients at the rotations
    # last, cur, next
    assert info.ncommit == pk.ncommit == 1 + 30 + 3 + 3 + 3 + 2 + 1 + 3 + 3
    assert info.z_ok == 1, "grand products that did not close: %d" % info.z_fail
    assert info.quotient_ok == 1
    assert info.open_ok == 1
    # h; instance, 30 advice, 4 fixed, random, 6 sigmas at x; permutation z at
    # x, xw (and x w^-(b+1) but the last); per lookup m, z, z(xw); shuffle z, z(xw)
    assert info.nevals == 1 + 1 + 30 + 4 + 1 + 6 + (3 * 3 - 1) + 3 * 3 + 2 * 2
    assert all(zo.g1a_is_on_curve(com[i:i + 1]) for i in range(len(com)))
    assert not (com == 0).all(axis=1).any()
    assert not (pieces == 0).all()


def test_prover_core_tamper_detected(core):
    from zkprover.prover import BATCHED, TAMPER_GATE
    pk, pr = core
    _, _, info = pr.prove(12, flags=BATCHED | TAMPER_GATE)
    assert info.z_ok == 1          # the gate is in no grand product
    assert info.quotient_ok == 0   # but h is no longer a polynomial


def test_prover_core_shplonk(core, zo):
    from zkprover.prover import BATCHED, SHPLONK
    pk, pr = core
    g, _, gi = pr.prove(13, flags=BATCHED)
    s, _, si = pr.prove(13, flags=BATCHED | SHPLONK)
    assert si.open_ok == 1 and si.quotient_ok == 1 and si.z_ok == 1
    assert si.ncommit == gi.ncommit - 1  # h(X) and L(X) against three GWC quotients
    base = gi.ncommit - 3
    assert (s[:base] == g[:base]).all()  # the phases before the opening are shared
    assert all(zo.g1a_is_on_curve(s[i:i + 1]) for i in range(base, si.ncommit))


@pytest.mark.parametrize("shplonk", [False, True])
def test_prover_core_opening_tamper_detected(core, shplonk):
    from zkprover.prover import BATCHED, SHPLONK, TAMPER_EVAL
    pk, pr = core
    _, _, info = pr.prove(14, flags=BATCHED | TAMPER_EVAL | (SHPLONK if shplonk else 0))
    assert info.quotient_ok == 1 and info.open_ok == 0


P_MOD = 0x30644e72e131a029b85045b68181585d97816a916871ca8d3c208c16d87cfd47
R_MOD = 21888242871839275222246405745257275088548364400416034343698204186575808495617


class _Reader:
    """a verifier reading a proof's bytes: points decompressed (x with the
    parity of y in the top bit; y^2 = x^3 + 3, p = 3 mod 4), scalars
    canonical; every item absorbed as halo2's Blake2bRead does (a model
    written with hashlib)"""

    def __init__(self, proof):
        import hashlib
        self.buf, self.pos = proof, 0
        self.h = hashlib.blake2b(digest_size=64, person=b"Halo2-Transcript")

    def common_point(self, x, y):
        self.h.update(b"\x01" + x.to_bytes(32, "little") + y.to_bytes(32, "little"))

    def common_scalar(self, s):
        self.h.update(b"\x02" + s.to_bytes(32, "little"))

    def read_point(self):
        b = self.buf[self.pos:self.pos + 32]
        self.pos += 32
        if b == bytes(32):
            self.common_point(0, 0)
            return (0, 0)
        odd = b[31] >> 7
        x = int.from_bytes(b[:31] + bytes([b[31] & 0x7f]), "little")
        y = pow((x * x * x + 3) % P_MOD, (P_MOD + 1) // 4, P_MOD)
        assert y * y % P_MOD == (x * x * x + 3) % P_MOD, "not a curve point"
        if y & 1 != odd:
            y = P_MOD - y
        self.common_point(x, y)
        return (x, y)

    def read_scalar(self):
        s = int.from_bytes(self.buf[self.pos:self.pos + 32], "little")
        self.pos += 32
        assert s < R_MOD
        self.common_scalar(s)
        return s

    def challenge(self):
        self.h.update(b"\x00")
        return int.from_bytes(self.h.copy().digest(), "little") % R_MOD


def _verify_transcript(zo, pk, com, proof, info, open_ch, shplonk):
    """reads the proof as the verifier does (src/lib.rs:242-1194 order): the
    key's representation and the instance commitments absorbed, then every
    written commitment and evaluation; returns the challenges it squeezes
    and the points it read"""
    c = pk.circuit
    A, L, P, S, I = c.advice, c.lookups, c.perm_cols, c.shuffles, c.instance
    Sp = (P + 1) // 2
    rd = _Reader(proof)
    rd.common_scalar(zo.from_mont(pk.vk_repr())[0])
    for i in range(I):
        rd.common_point(*zo.g1a_to_ints(com[i:i + 1]))
    pts = [rd.read_point() for _ in range(A)]
    rd.challenge()  # theta
    pts += [rd.read_point() for _ in range(L)]
    beta, gamma = rd.challenge(), rd.challenge()
    pts += [rd.read_point() for _ in range(Sp + L + S + 1)]
    y = rd.challenge()
    pts += [rd.read_point() for _ in range(3)]
    x = rd.challenge()
    nev = info.nevals - 1  # h(x) is not written
    [rd.read_scalar() for _ in range(nev)]
    ch = [beta, gamma, y, x]
    if shplonk:
        ch += [rd.challenge(), rd.challenge()]  # y', v
        pts.append(rd.read_point())
        ch.append(rd.challenge())  # u
        pts.append(rd.read_point())
    else:
        ch.append(rd.challenge())  # v
        while rd.pos < len(proof):
            pts.append(rd.read_point())
    assert rd.pos == len(proof)
    return ch, pts


@pytest.mark.parametrize("shplonk", [False, True])
def test_prover_core_proof_bytes_transcript(core, zo, shplonk):
    """the proof's bytes are a halo2 transcript: a verifier holding only the
    key's representation and the instance commitments reads every written
    commitment and evaluation from them and squeezes the same challenges
    (beta, gamma, y, x and the opening's) as the prover; the points it reads
    are the prover's commitments"""
    from zkprover.prover import BATCHED, SHPLONK
    pk, pr = core
    com, _, info = pr.prove(17, flags=BATCHED | (SHPLONK if shplonk else 0))
    proof = pr.proof_bytes()
    ev_m, och = pr.openings()
    ch, pts = _verify_transcript(zo, pk, com, proof, info, och, shplonk)
    got = [zo.from_mont(np.array(info.challenges[c], dtype=np.uint64).reshape(1, 4))[0] for c in range(4)]
    assert ch[:4] == got
    oc = zo.from_mont(och)
    if shplonk:
        assert ch[4:] == [oc[1], oc[0], oc[2]]  # y', v, u
    else:
        assert ch[4] == oc[0]
    I = pk.circuit.instance
    exp = [tuple(zo.g1a_to_ints(com[i:i + 1])) for i in range(I, info.ncommit)]
    assert [tuple(p) for p in pts] == exp
    # the written evaluations are the openings but h(x), in order: they follow
    # the advice, m, grand-product, random and quotient-piece points
    c = pk.circuit
    at = 32 * (c.advice + 2 * c.lookups + (c.perm_cols + 1) // 2 + c.shuffles + 1 + 3)
    got_ev = [int.from_bytes(proof[at + 32 * i:at + 32 * i + 32], "little") for i in range(info.nevals - 1)]
    assert got_ev == [int(v) for v in zo.from_mont(ev_m)[1:]]
    assert len(proof) == 32 * (info.ncommit - I) + 32 * (info.nevals - 1)


def test_prover_core_vk_repr(core, zo):
    """the key's transcript representation: Blake2b-512 personalised
    "Halo2-Verify-Key" of its canonical description, read mod r"""
    pk, _ = core
    rep = zo.from_mont(pk.vk_repr())[0]
    assert 0 < rep < R_MOD
    # the description covers the fixed and sigma commitments: a different key
    # seed gives a different representation
    from zkprover.prover import ProvingKey
    other = ProvingKey(pk.circuit, seed=4)
    try:
        assert zo.from_mont(other.vk_repr())[0] != rep
    finally:
        other.close()


def test_prover_core_batched_equals_plain_and_deterministic(core):
    from zkprover.prover import BATCHED
    pk, pr = core
    a = pr.prove(21, flags=BATCHED)
    b = pr.prove(21, flags=0)
    c = pr.prove(22, flags=BATCHED)
    assert (a[0] == b[0]).all() and (a[1] == b[1]).all()
    assert bytes(a[2].transcript) == bytes(b[2].transcript)
    assert bytes(a[2].transcript) != bytes(c[2].transcript)
    assert a[2].quotient_ok and b[2].quotient_ok and c[2].quotient_ok
    assert not (a[0] == c[0]).all()


def test_prover_core_k14(gpu):
    """a larger shape with the benchmark's column counts"""
    from zkprover.prover import Circuit, Prover, ProvingKey
    circ = Circuit.benchmark(14)
    pk = ProvingKey(circ, seed=5)
    pr = Prover(pk)
    try:
        _, _, info = pr.prove(31)
        assert info.z_ok == 1 and info.quotient_ok == 1 and info.open_ok == 1
        from zkprover.prover import BATCHED, SHPLONK
        _, _, info = pr.prove(32, flags=BATCHED | SHPLONK)
        assert info.z_ok == 1 and info.quotient_ok == 1 and info.open_ok == 1
    finally:
        pr.close()
        pk.close()


def test_prover_core_host_witness_same_proof(core):
    """a proof whose advice and instance columns come from host memory
    (ZKP_HOST_WITNESS, uploaded during the commitment phase) equals the proof
    that synthesised the same witness on the device: same commitments,
    quotient and transcript"""
    from zkprover import hostmem
    from zkprover.prover import BATCHED, HOST_WITNESS
    pk, pr = core
    n = 1 << pk.circuit.k
    c1, p1, i1 = pr.prove(23, flags=BATCHED)
    # the advice columns as prepare_advice_buffer hands them out: zeroed,
    # pinned; the instance column from the pool directly
    advb = hostmem.AdviceBuffer(n, pk.circuit.advice)
    assert all((advb.column(c) == 0).all() for c in range(pk.circuit.advice))
    assert len(set(advb.ptrs)) == pk.circuit.advice
    bufs = [hostmem.PinnedBuffer(n * 32) for _ in range(pk.circuit.instance)]
    try:
        adv = list(advb.ptrs)
        ins = [b.ptr for b in bufs]
        pr.witness_to_host(adv, ins)
        pr.set_host_witness(adv, ins)
        c2, p2, i2 = pr.prove(23, flags=BATCHED | HOST_WITNESS)
        assert i2.z_ok and i2.quotient_ok and i2.open_ok
        assert (c1 == c2).all() and (p1 == p2).all()
        assert bytes(i1.transcript) == bytes(i2.transcript)
        # pageable host memory works too (synchronous uploads)
        host = [advb.column(c).copy() for c in range(pk.circuit.advice)]
        host += [b.array(np.uint64).reshape(n, 4).copy() for b in bufs]
        pr.set_host_witness([h.ctypes.data for h in host[:pk.circuit.advice]],
                            [h.ctypes.data for h in host[pk.circuit.advice:]])
        c3, _, i3 = pr.prove(23, flags=BATCHED | HOST_WITNESS)
        assert (c1 == c3).all() and bytes(i1.transcript) == bytes(i3.transcript)
    finally:
        advb.close()
        for b in bufs:
            b.close()


def test_prover_core_graph_replays_equal_eager(gpu):
    """the h(X) round captured into a HIP graph (ZKP_GRAPH) on a job's first
    proof and replayed on the next ones, with eager proofs of other seeds and
    host synchronisations in between: every replayed proof equals the eager
    proof of the same seed (commitments, quotient, transcript)"""
    from zkprover.prover import BATCHED, GRAPH, SHPLONK, Prover, ProvingKey
    pk = ProvingKey(_small(11), seed=5)
    eager, graph = Prover(pk), Prover(pk)
    try:
        for seed, extra in ((61, 0), (62, SHPLONK), (63, 0), (64, 0)):
            ce, pe, ie = eager.prove(seed, flags=BATCHED | extra)
            cg, pg, ig = graph.prove(seed, flags=BATCHED | GRAPH | extra)
            assert ig.z_ok and ig.quotient_ok and ig.open_ok, seed
            assert (ce == cg).all() and (pe == pg).all(), seed
            assert bytes(ie.transcript) == bytes(ig.transcript), seed
            eager.prove(seed + 100, flags=BATCHED)  # other work between the replays
    finally:
        eager.close()
        graph.close()
        pk.close()
