"""GPU: the proof server admits requests of mixed sizes while others run,
keeps the per-proof device states within the HBM budget, and every result is
the same alone and in the mix (byte-identical commitments)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _shapes():
    from zkprover.prover import Circuit
    a = Circuit(k=10, advice=30, fixed=2, instance=1, lookups=3, table_log=7, perm_cols=6, shuffles=2,
                shuffle_log=7, blinding=5)
    b = Circuit(k=11, advice=30, fixed=2, instance=1, lookups=3, table_log=8, perm_cols=6, shuffles=2,
                shuffle_log=8, blinding=5)
    return [a, b]


@pytest.mark.parametrize("mode", ["eager", "cross_batch"])
def test_server_mix_equals_alone(gpu, mode):
    from zkprover.prover import BATCHED, footprint
    from zkprover.server import CROSS_BATCH, EAGER, Server
    shapes = _shapes()
    flags = BATCHED | EAGER | (CROSS_BATCH if mode == "cross_batch" else 0)
    plan = [(0, 101), (1, 102), (0, 103), (1, 104), (1, 105), (0, 106), (0, 107), (1, 108)]
    # alone: one slot, one request at a time
    alone = {}
    s1 = Server(shapes, slots=1)
    try:
        for sh, seed in plan:
            com, res = s1.wait(s1.submit(sh, seed), sh)
            assert res.err == 0 and res.z_ok and res.quotient_ok
            alone[seed] = com
    finally:
        s1.close()
    # in the mix: 3 slots, everything submitted at once, budget for 2 of the
    # larger per-proof states
    job_b = footprint(shapes[1])[1]
    s3 = Server(shapes, slots=3, hbm_budget=2 * job_b + job_b // 2, flags=flags)
    try:
        ids = [(s3.submit(sh, seed), sh, seed) for sh, seed in plan]
        for rid, sh, seed in ids:
            com, res = s3.wait(rid, sh)
            assert res.err == 0 and res.z_ok and res.quotient_ok
            assert (com == alone[seed]).all(), seed
        st = s3.stats()
        assert st["peak_running"] >= 2          # work overlapped
        assert st["peak_bytes"] <= st["budget_bytes"]
        bs = s3.batch_stats()
        if mode == "cross_batch":
            # every commitment group went through the batcher, and some
            # passes held groups of several proofs (all 8 were queued at once)
            assert bs["msm_items"] == sum(s3.ncommit[sh] for sh, _ in plan)
            assert bs["passes_with_several_proofs"] >= 1
        else:
            assert bs["msm_items"] == 0
    finally:
        s3.close()


def test_server_poisson_driver(gpu):
    from zkprover.server import Server, run_poisson, summarize
    s = Server(_shapes(), slots=2)
    try:
        recs, wall = run_poisson(s, rate_per_s=50.0, count=12, weights=[0.6, 0.4], seed=3)
        summ = summarize(recs, wall)
        assert summ["all_ok"] and summ["proofs"] == 12 and summ["p99_ms"] >= summ["p50_ms"] > 0
    finally:
        s.close()


def test_server_zkwasm_sizes_mix_equals_alone(gpu):
    """the serving workload's circuit family at k = 18, 20 and 22: every
    proof's commitments are the same run alone (one slot) and admitted into
    a mix of three slots while others are on the device, their commitment
    MSMs merged across proofs"""
    from zkprover.prover import Circuit
    from zkprover.server import Server
    shapes = [Circuit.benchmark(k) for k in (18, 20, 22)]
    plan = [(0, 201), (1, 202), (2, 203), (0, 204), (1, 205), (0, 206)]
    alone = {}
    s1 = Server(shapes, slots=1)
    try:
        for sh, seed in plan:
            com, res = s1.wait(s1.submit(sh, seed), sh)
            assert res.err == 0 and res.z_ok and res.quotient_ok, (sh, seed)
            alone[seed] = com
    finally:
        s1.close()
    from zkprover.prover import BATCHED
    from zkprover.server import CROSS_BATCH
    s3 = Server(shapes, slots=3, flags=BATCHED | CROSS_BATCH)
    try:
        ids = [(s3.submit(sh, seed), sh, seed) for sh, seed in plan]
        for rid, sh, seed in ids:
            com, res = s3.wait(rid, sh)
            assert res.err == 0 and res.z_ok and res.quotient_ok, (sh, seed)
            assert (com == alone[seed]).all(), seed
        assert s3.stats()["peak_running"] >= 2
    finally:
        s3.close()
