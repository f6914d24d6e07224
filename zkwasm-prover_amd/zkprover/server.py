"""Host binding of the proof server (zks_*) and a Poisson load driver.

The server is native (server.cpp): resident keys per circuit shape, worker
threads with their own streams, admission by per-proof HBM footprint.  This
module marshals the calls and, for measurement, replays an arrival process:
requests are submitted at exponential inter-arrival times and their
latencies (submit -> done) are collected.
"""
import ctypes
import time

import numpy as np

from ._lib import check, lib
from .prover import BATCHED, Circuit


class Result(ctypes.Structure):
    _fields_ = [("shape", ctypes.c_int), ("err", ctypes.c_int), ("z_ok", ctypes.c_int),
                ("quotient_ok", ctypes.c_int), ("ncommit", ctypes.c_int), ("reserved", ctypes.c_int),
                ("queue_ms", ctypes.c_double), ("run_ms", ctypes.c_double), ("latency_ms", ctypes.c_double),
                ("prove_ms", ctypes.c_double)]


def _decl():
    L = lib()
    L.zks_create.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_size_t, ctypes.c_int,
                             ctypes.c_ulonglong, ctypes.c_void_p]
    L.zks_submit.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_ulonglong]
    L.zks_submit.restype = ctypes.c_int
    L.zks_wait.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    L.zks_stats.argtypes = [ctypes.c_void_p] * 4
    L.zks_batch_stats.argtypes = [ctypes.c_void_p] * 4
    L.zks_destroy.argtypes = [ctypes.c_void_p]
    L.zkp_commitment_count.restype = ctypes.c_int
    L.zkp_commitment_count.argtypes = [ctypes.c_void_p]
    return L


EAGER = 0x100        # ZKS_EAGER: overlap whenever a slot and memory allow
CROSS_BATCH = 0x200  # ZKS_CROSS_BATCH: merge concurrent proofs' commitment MSMs


class Server:
    def __init__(self, circuits, slots=3, hbm_budget=0, flags=BATCHED, key_seed=1):
        self.L = _decl()
        self.circuits = list(circuits)
        arr = (Circuit * len(self.circuits))(*self.circuits)
        self.handle = ctypes.c_void_p()
        check(self.L.zks_create(arr, len(self.circuits), slots, hbm_budget, flags, key_seed,
                                ctypes.byref(self.handle)), "zks_create")
        self.ncommit = [self.L.zkp_commitment_count(ctypes.byref(c)) for c in self.circuits]

    def submit(self, shape, seed):
        rid = self.L.zks_submit(self.handle, shape, seed)
        if rid < 0:
            raise ValueError("zks_submit rejected the request")
        return rid

    def wait(self, rid, shape):
        com = np.zeros((self.ncommit[shape], 8), dtype=np.uint64)
        res = Result()
        check(self.L.zks_wait(self.handle, rid, com.ctypes.data, ctypes.byref(res)), "zks_wait")
        return com, res

    def stats(self):
        pr, pb, bud = ctypes.c_int(), ctypes.c_size_t(), ctypes.c_size_t()
        check(self.L.zks_stats(self.handle, ctypes.byref(pr), ctypes.byref(pb), ctypes.byref(bud)), "zks_stats")
        return {"peak_running": pr.value, "peak_bytes": pb.value, "budget_bytes": bud.value}

    def batch_stats(self):
        l, m, i = ctypes.c_ulonglong(), ctypes.c_ulonglong(), ctypes.c_ulonglong()
        check(self.L.zks_batch_stats(self.handle, ctypes.byref(l), ctypes.byref(m), ctypes.byref(i)),
              "zks_batch_stats")
        return {"merged_passes": l.value, "passes_with_several_proofs": m.value, "msm_items": i.value}

    def close(self):
        if self.handle:
            check(self.L.zks_destroy(self.handle), "zks_destroy")
            self.handle = ctypes.c_void_p()


def run_poisson(server, rate_per_s, count, weights, seed=10):
    """submits `count` requests with exponential inter-arrival times of mean
    1/rate and shapes drawn with `weights`; returns per-request records and
    the wall time from the first submit to the last completion"""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, count)
    shapes = rng.choice(len(weights), size=count, p=np.asarray(weights) / np.sum(weights))
    ids = []
    t0 = time.perf_counter()
    due = t0
    for i in range(count):
        due += gaps[i]
        delay = due - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        ids.append((server.submit(int(shapes[i]), 1000 + i), int(shapes[i])))
    recs = []
    for rid, sh in ids:
        _, res = server.wait(rid, sh)
        recs.append({"shape": sh, "latency_ms": res.latency_ms, "queue_ms": res.queue_ms, "run_ms": res.run_ms,
                     "ok": bool(res.z_ok and res.quotient_ok and res.err == 0)})
    wall = time.perf_counter() - t0
    return recs, wall


def summarize(recs, wall_s):
    lat = np.array([r["latency_ms"] for r in recs])
    return {"proofs": len(recs), "wall_s": round(wall_s, 3),
            "proofs_per_hour": round(len(recs) / wall_s * 3600.0, 1),
            "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
            "all_ok": all(r["ok"] for r in recs)}
