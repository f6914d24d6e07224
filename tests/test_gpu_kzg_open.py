"""GPU: the prover core's openings verify as KZG openings.

The key's SRS is a KZG setup for a seeded secret tau (g[i] = tau^i G,
g_lag[i] = L_i(tau) G), so every commitment is [p(tau)] G.  With the
trapdoor the verifier's pairing equations become equations in G1, checked
with the oracle's curve arithmetic (tests/verifier_model.py restates the
verifier: the reference's evaluation and query order, GWC groups by rotation,
Shplonk rotation sets):
  * the two bases agree: committing coefficients against g and the same
    polynomial's values against g_lag give one point (for the fixed columns
    of the verifying key, rebuilt from the key's own SRS);
  * GWC (src/multiopen.rs:16-177): for each rotation group with point z, the
    batch C = sum v^(len-1-i) C_i and E = sum v^(len-1-i) e_i over its
    queries, (tau - z) W_z = C - E G;
  * Shplonk (src/multiopen.rs:178-561): (tau - u) [L] =
    (sum_s v^(k-1-s) z_s (V_s - r_s(u) G) - Z_T(u) [h]) / z_0.
A wrong evaluation, quotient or challenge breaks the equation for all but a
negligible set of tau.  The same openings also pass the verifier's own
pairing checks, which hold only [tau]_2 = tau G2 of the setup (the pairing of
oracle/bn254_pairing.py): GWC as one product check over all points with a
random u, Shplonk as e([L], [tau]_2) = e(u [L] + F / z_0, G2); the tampered
proof fails it.  The vanishing argument is checked the verifier's way
too: h(x) (x^n - 1) equals the composed constraint rebuilt from the opened
evaluations and the challenges.
"""
import numpy as np
import pytest

import verifier_model as vm

pytestmark = pytest.mark.gpu

R = vm.R


def _circuit():
    from zkprover.prover import Circuit
    return Circuit(k=10, advice=30, fixed=4, instance=1, lookups=3, table_log=7, perm_cols=6, shuffles=2,
                   shuffle_log=7, blinding=5)


@pytest.fixture(scope="module")
def setup(gpu):
    from zkprover.prover import Prover, ProvingKey
    pk = ProvingKey(_circuit(), seed=9)
    pr = Prover(pk)
    yield pk, pr
    pr.close()
    pk.close()


def _proof(pr, zo, seed, flags):
    com, _, info = pr.prove(seed, flags=flags)
    ev_m, och = pr.openings()
    return com, info, vm.ints(zo, ev_m), vm.challenges(zo, info, och)


def test_fixed_commitments_agree_across_bases(setup, zo):
    """the verifying key (coefficients against g) equals the Lagrange-basis
    commitment of the same columns' values: g and g_lag come from one tau"""
    pk, _ = setup
    tau = vm.ints(zo, pk.srs_secret())[0]
    vk = pk.fixed_commitments()
    assert len(vk) == pk.circuit.fixed + pk.circuit.perm_cols
    assert all(zo.g1a_is_on_curve(vk[i:i + 1]) for i in range(len(vk)))
    # q is l_active: 1 on rows [0, u); its polynomial at tau is sum_{i<u} L_i(tau)
    circ = pk.circuit
    n, u = 1 << circ.k, (1 << circ.k) - (circ.blinding + 1)
    w = vm.ints(zo, zo.root_of_unity(circ.k))[0]
    c = (pow(tau, n, R) - 1) * pow(n, -1, R) % R
    q_tau = sum(pow(w, i, R) * pow(tau - pow(w, i, R), -1, R) for i in range(u)) * c % R
    g1 = vm.G1(zo)
    assert (g1.aff(g1.mul_aff(g1.g, q_tau)) == vk[0:1]).all()


def test_instance_commitment_is_the_verifiers(setup, zo):
    """the instance commitment the proof absorbs is the one a verifier
    computes from the public values alone: sum_i inst_i L_i(tau) G"""
    from zkprover.prover import BATCHED
    pk, pr = setup
    com, _, info = pr.prove(40, flags=BATCHED)
    inst = pr.instance()
    circ = pk.circuit
    assert inst.shape == (1, 1 << circ.k, 4)
    tau = vm.ints(zo, pk.srs_secret())[0]
    n, u = 1 << circ.k, (1 << circ.k) - (circ.blinding + 1)
    w = vm.ints(zo, zo.root_of_unity(circ.k))[0]
    vals = vm.ints(zo, inst[0])
    assert all(v == 0 for v in vals[u:])  # padded with zeros past the usable rows
    c = (pow(tau, n, R) - 1) * pow(n, -1, R) % R
    acc = sum(vals[i] * pow(w, i, R) % R * pow(tau - pow(w, i, R), -1, R) for i in range(n) if vals[i]) * c % R
    g1 = vm.G1(zo)
    assert (g1.aff(g1.mul_aff(g1.g, acc)) == com[0:1]).all()


@pytest.mark.parametrize("seed", [41, 42])
def test_gwc_openings_verify(setup, zo, seed):
    from zkprover.prover import BATCHED
    pk, pr = setup
    com, info, ev, ch = _proof(pr, zo, seed, BATCHED)
    assert info.open_ok == 1
    tau = vm.ints(zo, pk.srs_secret())[0]
    assert vm.gwc_accepts(pk.circuit, com, pk.fixed_commitments(), ev, ch, tau, zo, vm.G1(zo),
                          pairing=seed == 41)


def test_shplonk_opening_verifies(setup, zo):
    from zkprover.prover import BATCHED, SHPLONK
    pk, pr = setup
    com, info, ev, ch = _proof(pr, zo, 43, BATCHED | SHPLONK)
    assert info.open_ok == 1
    tau = vm.ints(zo, pk.srs_secret())[0]
    assert vm.shplonk_accepts(pk.circuit, com, pk.fixed_commitments(), ev, ch, tau, zo, vm.G1(zo))


def test_tampered_evaluation_fails_verification(setup, zo):
    """one evaluation off by one: the GWC equation at its point must fail"""
    from zkprover.prover import BATCHED, TAMPER_EVAL
    pk, pr = setup
    com, info, ev, ch = _proof(pr, zo, 44, BATCHED | TAMPER_EVAL)
    tau = vm.ints(zo, pk.srs_secret())[0]
    g1 = vm.G1(zo)
    groups = vm.gwc_groups(pk.circuit, com, pk.fixed_commitments(), ev, ch, zo, g1)
    ok = [bool((g1.aff(g1.mul_aff(w, tau - z)) == g1.aff(rhs)).all()) for w, z, rhs in groups]
    assert not all(ok)
    assert not vm.gwc_pairing_accepts(groups, tau, g1)


@pytest.mark.parametrize("tamper", [False, True])
def test_quotient_identity_at_x(setup, zo, tamper):
    """the vanishing argument: h(x) (x^n - 1) = C(x) from the opened values
    (h(x) is the folded quotient's evaluation); a witness with one broken
    gate row makes h no polynomial, and the identity fails"""
    from zkprover.prover import BATCHED, TAMPER_GATE
    pk, pr = setup
    com, info, ev, ch = _proof(pr, zo, 45, BATCHED | (TAMPER_GATE if tamper else 0))
    ok = vm.quotient_identity_holds(pk.circuit, com, pk.fixed_commitments(), ev, ch, zo, vm.G1(zo))
    assert ok == (not tamper)
