"""Autotuning of the MSM window and the NTT radix plan per size class.

The timing loops are native (zk_msm_autotune, zk_ntt_autotune in the kernel
library): each times the candidates on synthetic device data and records the
fastest in the library's per-size table, which every later `msm`, `zk_msm`,
`zk_msm_batch`, `ntt` and `zk_ntt_batch` call reads.  This module generates
the synthetic inputs, runs the tuners over a set of sizes and saves / restores
the table as JSON so a serving process can start from a tuned state.

The reference picks its window with a fixed formula (src/cuda/msm.rs:105, :275)
and its NTT radix with a fixed max_deg (src/cuda/ntt.rs:61-93); both become
measured choices here.
"""
import ctypes
import json

import torch

from . import device as dv
from ._lib import call, lib

_R = 21888242871839275222246405745257275088548364400416034343698204186575808495617


def _root_of_unity(log_n):
    w = pow(7, (_R - 1) >> 28, _R)
    for _ in range(log_n, 28):
        w = w * w % _R
    return w


def _mont_limbs(v):
    m = (v << 256) % _R
    return [(m >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(4)]


def omega_table(log_n, inverse=False):
    """the w^i table of `expand_omega_buffer` for 2^log_n (device tensor)"""
    import numpy as np
    w = _root_of_unity(log_n)
    if inverse:
        w = pow(w, -1, _R)
    init = np.zeros((1 << log_n, 4), dtype=np.uint64)
    init[0] = _mont_limbs(1)
    if log_n:
        init[1] = _mont_limbs(w)
    t = dv.to_dev(init)
    call("expand_omega_buffer", t.data_ptr(), log_n, 1)
    torch.cuda.synchronize()
    return t


def tune_msm(log_sizes, reps=2, seed=31):
    """runs the MSM tuner for 2^k points for each k; returns {k: (c, {c: ms})}"""
    out = {}
    st = dv.stream_handle()
    nmax = 1 << max(log_sizes)
    pts = torch.empty(nmax * 64, dtype=torch.uint8, device=dv.device())
    sc = torch.empty(nmax * 32, dtype=torch.uint8, device=dv.device())
    call("zk_gen_points", pts.data_ptr(), nmax, seed, st)
    call("zk_rand_fr", sc.data_ptr(), nmax, seed + 1, st)
    torch.cuda.synchronize()
    for k in log_sizes:
        best = ctypes.c_int()
        ms = (ctypes.c_float * 18)()
        call("zk_msm_autotune", pts.data_ptr(), sc.data_ptr(), 1 << k, reps, st, ctypes.byref(best), ms)
        out[k] = (best.value, {c: round(ms[c], 4) for c in range(18) if ms[c] > 0})
    return out


def tune_ntt(log_sizes, reps=5, seed=41):
    """runs the NTT radix tuner for each 2^k; returns {k: (max_deg, {deg: ms})}"""
    out = {}
    st = dv.stream_handle()
    for k in log_sizes:
        n = 1 << k
        om = omega_table(k)
        a = torch.empty(n * 32, dtype=torch.uint8, device=dv.device())
        b = torch.empty_like(a)
        call("zk_rand_fr", a.data_ptr(), n, seed + k, st)
        best = ctypes.c_int()
        ms = (ctypes.c_float * 9)()
        call("zk_ntt_autotune", a.data_ptr(), b.data_ptr(), om.data_ptr(), k, reps, st, ctypes.byref(best), ms)
        out[k] = (best.value, {d: round(ms[d], 4) for d in range(9) if ms[d] > 0})
        del om, a, b
    return out


def table():
    """the library's current per-size choices"""
    L = lib()
    L.zk_msm_get_window.restype = ctypes.c_int
    L.zk_ntt_get_max_deg.restype = ctypes.c_int
    return {"msm_window": {k: L.zk_msm_get_window(k) for k in range(1, 31)},
            "ntt_max_deg": {k: L.zk_ntt_get_max_deg(k) for k in range(1, 31)}}


def save(path, msm=None, ntt=None):
    """writes the tuned entries (from tune_msm / tune_ntt results) as JSON"""
    data = {"msm_window": {str(k): v[0] for k, v in (msm or {}).items()},
            "ntt_max_deg": {str(k): v[0] for k, v in (ntt or {}).items()}}
    with open(path, "w") as f:
        json.dump(data, f, indent=1)


def load(path):
    """restores a saved table into the library"""
    with open(path) as f:
        data = json.load(f)
    for k, c in data.get("msm_window", {}).items():
        call("zk_msm_set_window", int(k), int(c))
    for k, d in data.get("ntt_max_deg", {}).items():
        call("zk_ntt_set_max_deg", int(k), int(d))
