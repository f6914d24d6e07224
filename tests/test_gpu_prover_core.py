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
    return Circuit(k=k, advice=30, fixed=2, instance=1, lookups=3, table_log=7, perm_cols=6, shuffles=2,
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
    # advice, m, logup z, permutation z (3 sets), shuffle z, quotient pieces,
    # GWC quotients at x, x w, x w^u
    assert info.ncommit == pk.ncommit == 30 + 2 * 3 + 3 + 2 + 3 + 3
    assert info.z_ok == 1, "grand products that did not close: %d" % info.z_fail
    assert info.quotient_ok == 1
    assert info.open_ok == 1
    # set {x}: 30 advice, q, t, 6 sigmas, 3 m, h; {x, xw}: 3 logup z, 2
    # shuffle z, last permutation z; {x, xw, xw^u}: 2 permutation z
    assert info.nevals == 42 + 2 * 6 + 3 * 2
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


def _transcript_model(zo, com, A, L, nz):
    """beta, gamma, y, x as halo2's Blake2bWrite derives them from the
    commitments (a model written with hashlib): advice, theta, m, beta,
    gamma, the grand products, y, the quotient pieces, x"""
    import hashlib
    R = 21888242871839275222246405745257275088548364400416034343698204186575808495617
    h = hashlib.blake2b(digest_size=64, person=b"Halo2-Transcript")

    def point(i):
        x, y = zo.g1a_to_ints(com[i:i + 1])
        h.update(b"\x01" + x.to_bytes(32, "little") + y.to_bytes(32, "little"))

    def challenge():
        h.update(b"\x00")
        return int.from_bytes(h.copy().digest(), "little") % R

    i = 0
    for _ in range(A):
        point(i); i += 1
    challenge()  # theta
    for _ in range(L):
        point(i); i += 1
    beta, gamma = challenge(), challenge()
    for _ in range(nz):
        point(i); i += 1
    y = challenge()
    for _ in range(3):
        point(i); i += 1
    return [beta, gamma, y, challenge()]


def test_prover_core_transcript(core, zo):
    """the challenges are the transcript's: recomputed from the returned
    commitments with a hashlib model of halo2's Blake2bWrite"""
    from zkprover.prover import BATCHED, SHPLONK
    pk, pr = core
    com, _, info = pr.prove(17)
    exp = _transcript_model(zo, com, 30, 3, 3 + 3 + 2)
    got = [zo.from_mont(np.array(info.challenges[c], dtype=np.uint64).reshape(1, 4))[0] for c in range(4)]
    assert got == exp
    com2, _, info2 = pr.prove(17, flags=BATCHED | SHPLONK)
    assert bytes(info.transcript) != bytes(info2.transcript)  # the openings differ
    assert [list(info2.challenges[c]) for c in range(4)] == [list(info.challenges[c]) for c in range(4)]


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
