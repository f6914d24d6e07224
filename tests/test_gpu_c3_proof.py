"""GPU: a full proof of the benchmark circuit at k = 20 (40 advice, 10 fixed,
1 instance column; the gate a*b + c - d = 0 and 8 coefficient gates over the
fixed columns; 6 logup lookups into a 2^16 range table; 12 permutation
columns, the instance column among them, with random copy cycles; 2 shuffles
of 2^18 rows), opened with GWC and with Shplonk, checked the verifier's way
(tests/verifier_model.py):

  * every grand product closes and h(X) is a polynomial (degree < 3n);
  * the vanishing argument at x: h(x)(x^n - 1) equals the constraint
    composed from the opened evaluations and the transcript's challenges;
  * the GWC openings satisfy (tau - z) W_z = C_z - E_z G at every rotation
    and pass the verifier's pairing product check, holding only [tau]_2;
  * the Shplonk opening satisfies the verifier's pairing equation.
"""
import pytest

import verifier_model as vm

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c3(gpu):
    from zkprover.prover import Circuit, Prover, ProvingKey
    circ = Circuit.benchmark(20)
    assert (circ.advice, circ.fixed, circ.instance, circ.lookups, circ.perm_cols, circ.shuffles) == \
        (40, 10, 1, 6, 12, 2) and circ.shuffle_log == 18
    pk = ProvingKey(circ, seed=20)
    pr = Prover(pk)
    yield pk, pr
    pr.close()
    pk.close()


@pytest.mark.parametrize("shplonk", [False, True])
def test_c3_k20_proof_verifies(c3, zo, shplonk):
    from zkprover.prover import BATCHED, SHPLONK
    pk, pr = c3
    com, _, info = pr.prove(2001 + shplonk, flags=BATCHED | (SHPLONK if shplonk else 0), want_pieces=False)
    assert info.z_ok == 1 and info.quotient_ok == 1 and info.open_ok == 1
    ev_m, och = pr.openings()
    ev, ch = vm.ints(zo, ev_m), vm.challenges(zo, info, och)
    vk = pk.fixed_commitments()
    assert len(vk) == 10 + 12
    g1 = vm.G1(zo)
    assert vm.quotient_identity_holds(pk.circuit, com, vk, ev, ch, zo, g1)
    tau = vm.ints(zo, pk.srs_secret())[0]
    if shplonk:
        assert vm.shplonk_accepts(pk.circuit, com, vk, ev, ch, tau, zo, g1)
    else:
        assert vm.gwc_accepts(pk.circuit, com, vk, ev, ch, tau, zo, g1)
