"""Benchmark of the halo2 KZG arithmetic core at zkWasm size on one MI355X.

One step = one pass over the primitive set, inputs already resident in HBM:
  G1 MSM n = 2^22 (msm + batch_msm_collect, the reference host-driver buffer
  policy), Fr NTT 2^22, iNTT 2^22, coset NTT 2^22 -> 2^24 (extended_prepare +
  NTT 2^24) and batch inversion of 2^22 nonzero Fr.
`value` is the 2^22 MSM time in ms (lower is better); every primitive's time
is reported beside it.  The roofline object covers the dominant kernel (the
MSM bucket accumulation), timed with HIP events on its own stream.
The CPU baseline is the C oracle on the host cores given to the job: every
primitive once at n = 2^22 (the MSM is its `value`) and a k=20 proof lower
bound.

Beside it (the metric's other two terms): 64 uneven MSMs and 64 NTTs in one
batched pass against the plain per-call path; the prover core's k=20 and
k=22 proofs (median of 5 per mode, device and host-resident witness); and
the serving leg, plain one-at-a-time against the server at four offered
loads, p50 / p99 latency and proofs per hour.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batch] [--no-proof] [--no-serve]
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "zkwasm-prover_amd"))

METRIC = "k=22 proof time (s); 2^22 MSM ms; batched proofs/hr vs plain at p99"
LOG_N = 22
EXT = 2
# gfx950 VALU: 256 CUs x 64 lanes x 2.4 GHz, one 32-bit op per lane per cycle
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
HBM_PEAK_GBS = 8000.0
COPY_CAL_BYTES = 256 << 20
# canonical work of one point-to-bucket addition: the Jacobian mixed add
# (7M + 4S = 11 Fp multiplies), each a 8x8-limb CIOS product = 136 mad ops
MULS_PER_BUCKET_ADD = 11
OPS_PER_MONT_MUL = 136


def progress(msg):
    """one progress line on stderr (the JSON result is the only stdout line)"""
    print("bench: " + msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-proof", action="store_true")
    ap.add_argument("--proof-ks", default="20,22")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-count", type=int, default=32)
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from zkprover import device as dv
    from zkprover._lib import call, lib

    L = lib()
    L.zk_rand_fr.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_void_p]
    L.zk_gen_points.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_void_p]
    L.zk_msm_profile_enable.argtypes = [ctypes.c_int]
    L.zk_msm_profile_enable.restype = None
    L.zk_msm_profile_read.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.zk_msm_profile_read.restype = ctypes.c_int
    L.zk_msm_default_window.argtypes = [ctypes.c_int]
    L.zk_msm_default_window.restype = ctypes.c_int
    L.zkp_msm_policy.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 3
    L.zkp_msm_policy.restype = None

    dev = dv.device()
    st = dv.stream_handle()
    n = 1 << LOG_N
    N = 1 << (LOG_N + EXT)

    def u8(nbytes):
        return torch.empty(nbytes, dtype=torch.uint8, device=dev)

    # ---- inputs (seeded, generated on the device) -------------------------
    points = u8(n * 64)
    scalars = u8(n * 32)
    call("zk_gen_points", points.data_ptr(), n, 4, st)
    call("zk_rand_fr", scalars.data_ptr(), n, 3, st)
    ntt_in = u8(n * 32)
    call("zk_rand_fr", ntt_in.data_ptr(), n, 5, st)
    coset_in = u8(n * 32)
    call("zk_rand_fr", coset_in.data_ptr(), n, 6, st)
    inv_in = u8(n * 32)
    call("zk_rand_fr", inv_in.data_ptr(), n, 7, st)  # zero has probability ~2^-232

    # twiddle tables as src/cuda/ntt.rs:36-59 builds them
    def omega_table(log_n, inverse):
        w = _root_of_unity_int(log_n)
        if inverse:
            w = pow(w, -1, _R)
        init = np.zeros((1 << log_n, 4), dtype=np.uint64)
        init[0] = _to_mont_limbs(1)
        init[1] = _to_mont_limbs(w)
        t = dv.to_dev(init)
        call("expand_omega_buffer", t.data_ptr(), log_n, 1)
        return t

    om_n, om_n_inv, om_N = omega_table(LOG_N, False), omega_table(LOG_N, True), omega_table(LOG_N + EXT, False)
    n_inv = dv.to_dev(np.array([_to_mont_limbs(pow(n, -1, _R))], dtype=np.uint64))
    zeta = pow(7, (_R - 1) // 3, _R)
    coset_powers = dv.to_dev(np.array([_to_mont_limbs(zeta), _to_mont_limbs(zeta * zeta % _R)], dtype=np.uint64))

    # MSM buffers per the reference's host driver (src/cuda/msm.rs:284-311)
    wb, windows, workers = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    L.zkp_msm_policy(1, n, ctypes.byref(wb), ctypes.byref(windows), ctypes.byref(workers))
    curve_slice = u8((((windows.value << wb.value) + 2 * workers.value) * 96))
    single = n * windows.value
    idx = [u8(single * 4) for _ in range(4)]
    sort_buf = u8(single * 3 * 4)
    acc_idx = u8(workers.value * 2 * 4)
    slice_ptrs = dv.to_dev(np.array([curve_slice.data_ptr()], dtype=np.uint64).reshape(1, 1))

    buf_a, buf_b = u8(N * 32), u8(N * 32)
    inv_buf, inv_tmp = u8(n * 32), u8(n * 32)

    # lookup multiplicity: the benchmark circuit's range table (row j holds
    # j mod 2^16) and one input of uniform range values; buffers as the
    # reference's caller allocates them (src/lib.rs:407-412)
    range_vals = np.array([_to_mont_limbs(v) for v in range(1 << 16)], dtype=np.uint64)
    lk_table = dv.to_dev(range_vals[np.arange(n) % (1 << 16)])
    lk_input = dv.to_dev(range_vals[np.random.default_rng(5).integers(0, 1 << 16, size=n)])
    lk_m, lk_so, lk_t0, lk_t1 = u8(n * 32), u8(n * 32), u8(n * 32), u8(n * 32)
    lk_u = n - 6

    # one eval_h constraint kernel at the extended size of a k=22 proof: the
    # logup term over 2^24 rows (ten input columns, res read and written)
    eh = [u8(N * 32) for _ in range(10)]
    for i, t in enumerate(eh):
        call("zk_rand_fr", t.data_ptr(), N, 40 + i, st)
    eh_y = u8(32)
    call("zk_rand_fr", eh_y.data_ptr(), 1, 50, st)

    def run_logup_eval_h():
        call("logup_eval_h", *[t.data_ptr() for t in eh], eh_y.data_ptr(), 4, N, st)

    def run_prepare_table():
        call("prepare_table_lookup", lk_table.data_ptr(), None, None, None, None, lk_so.data_ptr(), None, 0,
             LOG_N, lk_u, st)

    def run_calc_m():
        call("calc_m", lk_m.data_ptr(), lk_table.data_ptr(), lk_input.data_ptr(), None, lk_so.data_ptr(),
             lk_t0.data_ptr(), lk_t1.data_ptr(), None, 0, LOG_N, lk_u, st)
    swap = ctypes.c_bool(False)

    def run_msm():
        call("msm", curve_slice.data_ptr(), points.data_ptr(), scalars.data_ptr(), idx[0].data_ptr(),
             idx[1].data_ptr(), idx[2].data_ptr(), idx[3].data_ptr(), sort_buf.data_ptr(), acc_idx.data_ptr(),
             n, windows.value, wb.value, 64, workers.value, min(single * 12, 0x7fffffff), 0, st)
        call("batch_msm_collect", None, None, None, None, slice_ptrs.data_ptr(), workers.value, windows.value,
             wb.value, 1, st)

    def run_ntt(log_n, om, ninv):
        call("ntt", buf_a.data_ptr(), buf_b.data_ptr(), None, om.data_ptr(), dv.ptr(ninv), log_n, 8,
             ctypes.addressof(swap), st)

    def run_coset():
        call("extended_prepare", buf_a.data_ptr(), coset_powers.data_ptr(), 3, n, N, 0, st)
        call("ntt", buf_a.data_ptr(), buf_b.data_ptr(), None, om_N.data_ptr(), None, LOG_N + EXT, 8,
             ctypes.addressof(swap), st)

    class _CosetItem(ctypes.Structure):
        _fields_ = [("src", ctypes.c_void_p), ("buf", ctypes.c_void_p), ("tmp", ctypes.c_void_p),
                    ("omegas", ctypes.c_void_p), ("log_n", ctypes.c_int), ("log_ext", ctypes.c_int)]
    L.zk_coset_extend_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_void_p]
    coset_item = _CosetItem(coset_in.data_ptr(), buf_a.data_ptr(), buf_b.data_ptr(), om_N.data_ptr(), LOG_N, EXT)
    coset_sw = ctypes.c_ubyte(0)

    def run_coset_fused():
        # the same extension in one pipeline: the first pass reads the 2^22
        # coefficients and scales them on load
        if L.zk_coset_extend_batch(ctypes.addressof(coset_item), 1, coset_powers.data_ptr(), 3,
                                   ctypes.addressof(coset_sw), st) != 0:
            raise RuntimeError("zk_coset_extend_batch")

    def restore(dst, src):
        return lambda: dst[: src.numel()].copy_(src)

    # (name, untimed input restore, timed primitive)
    prims = [("msm_2^22", None, run_msm),
             ("ntt_2^22", restore(buf_a, ntt_in), lambda: run_ntt(LOG_N, om_n, None)),
             ("intt_2^22", restore(buf_a, ntt_in), lambda: run_ntt(LOG_N, om_n_inv, n_inv)),
             ("coset_ntt_2^22_to_2^24", restore(buf_a, coset_in), run_coset),
             ("coset_extend_fused_2^22_to_2^24", None, run_coset_fused),
             ("batch_inversion_2^22", restore(inv_buf, inv_in),
              lambda: call("zk_batch_invert", inv_buf.data_ptr(), inv_tmp.data_ptr(), n, st)),
             ("prepare_table_lookup_2^22", None, run_prepare_table),
             ("calc_m_2^22", None, run_calc_m),
             ("logup_eval_h_2^24", None, run_logup_eval_h)]

    # one plain copy of 256 MiB: the calibration dispatch for the HBM byte
    # counters (FETCH_SIZE / WRITE_SIZE) when this script runs under rocprofv3
    call("zk_copy", buf_b.data_ptr(), buf_a.data_ptr(), COPY_CAL_BYTES, st)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        for _, prep, fn in prims:
            if prep:
                prep()
            fn()
    torch.cuda.synchronize()

    # ---- timed region --------------------------------------------------------
    L.zk_msm_profile_enable(1)
    times = np.zeros(8, dtype=np.float64)
    L.zk_msm_profile_read(times.ctypes.data, 8, 1)
    ev = {name: [] for name, _, _ in prims}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for name, prep, fn in prims:
            if prep:
                prep()
            s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record()
            fn()
            e.record()
            ev[name].append((s, e))
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3
    calls = L.zk_msm_profile_read(times.ctypes.data, 8, 1)
    L.zk_msm_profile_enable(0)
    prim_ms = {name: sum(s.elapsed_time(e) for s, e in lst) / len(lst) for name, lst in ev.items()}
    phase_names = ["digits_hist", "scan", "scatter", "accumulate", "fixup", "reduce", "header"]
    phases = {k: times[i] / max(calls, 1) for i, k in enumerate(phase_names)}
    step_ms = wall_ms / args.steps  # includes the untimed input restores (D2D copies)

    # ---- roofline of the dominant kernel (bucket accumulation) --------------
    c = L.zk_msm_default_window(n)
    w_used = math.ceil(255 / c)
    units = n * w_used  # point-to-bucket additions per launch (zero digits have probability ~2^-c)
    ops = units * MULS_PER_BUCKET_ADD * OPS_PER_MONT_MUL
    acc_ms = phases["accumulate"]
    achieved = ops / (acc_ms * 1e-3) / 1e12
    roofline = {
        "bound": "valu", "kernel": "k_msm_accumulate", "achieved": round(achieved, 3),
        "peak": round(VALU_PEAK_TOPS, 2), "unit": "TOP/s", "frac": round(achieved / VALU_PEAK_TOPS, 4),
        "traffic": _pmc_traffic("k_msm_accumulate"), "avg_launch_ms": round(acc_ms, 4), "window_bits": c,
        "algorithmic": f"{units} bucket additions x {MULS_PER_BUCKET_ADD} Fp mont-mul x "
                       f"{OPS_PER_MONT_MUL} 32-bit mad ops",
    }

    # the other hot kernels against both ceilings: algorithmic bytes (one read
    # of every input and one write of every output) over HBM peak, canonical
    # Montgomery products (136 mad ops each) over VALU peak
    def fracs(ms, nbytes, muls):
        return {"time_ms": round(ms, 4), "bytes": nbytes, "mont_muls": muls,
                "hbm_frac": round(nbytes / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
                "valu_frac": round(muls * OPS_PER_MONT_MUL / (ms * 1e-3) / (VALU_PEAK_TOPS * 1e12), 4)}
    secondary = {
        "ntt_2^22": fracs(prim_ms["ntt_2^22"], 2 * n * 32, n // 2 * LOG_N),
        "intt_2^22": fracs(prim_ms["intt_2^22"], 2 * n * 32, n // 2 * LOG_N + n),
        "coset_ntt_2^22_to_2^24": fracs(prim_ms["coset_ntt_2^22_to_2^24"], n * 32 + N * 32,
                                        N // 2 * (LOG_N + EXT) + n),
        "coset_extend_fused_2^22_to_2^24": fracs(prim_ms["coset_extend_fused_2^22_to_2^24"], n * 32 + N * 32,
                                                 N // 2 * (LOG_N + EXT) + n),
        "batch_inversion_2^22": fracs(prim_ms["batch_inversion_2^22"], 2 * n * 32, 3 * n),
        "logup_eval_h_2^24": fracs(prim_ms["logup_eval_h_2^24"], 11 * N * 32, 9 * N),
    }

    out = {
        "metric": METRIC, "value": round(prim_ms["msm_2^22"], 3), "unit": "ms (G1 MSM, n=2^22)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(step_ms, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32 limbs (exact 256-bit Montgomery)", "data": "synthetic",
        "config": {"workload": "zkWasm-size primitives: BN254 G1 MSM n=2^22; Fr NTT and iNTT 2^22; "
                               "coset NTT 2^22->2^24; batch inversion 2^22; lookup multiplicity "
                               "(range table 2^16 in 2^22 rows, one input); logup eval_h term over 2^24 rows",
                   "msm_window_bits": c,
                   "parallelism": "single GPU"},
        "primitives_ms": {k: round(v, 4) for k, v in prim_ms.items()},
        "msm_phases_ms": {k: round(v, 4) for k, v in phases.items()},
        "roofline": roofline,
        "secondary_rooflines": secondary,
    }
    progress("primitives done: MSM 2^22 %.3f ms" % prim_ms["msm_2^22"])
    if not args.no_batch:
        out["batched"] = {"msm": bench_batched_msm(L, points, max(1, args.steps // 2)),
                          "ntt": bench_batched_ntt(L, args.steps),
                          "msm_fixed_base_tables": bench_msm_table(L)}
        progress("batched done")
    if not args.no_proof:
        out["proof_core"] = bench_proof_core([int(k) for k in args.proof_ks.split(",") if k])
    if not args.no_serve:
        out["serving"] = bench_serving(args.serve_count)
    out["host_copy"] = bench_host_copy(prim_ms["msm_2^22"])
    if not args.no_cpu:
        progress("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(n)
    print(json.dumps(out), flush=True)


def bench_host_copy(msm_ms, reps=5):
    """Host <-> device rates of one k=22 column (2^27 bytes): pinned pool
    memory against pageable numpy memory, hipMemcpyAsync timed with HIP
    events on the copying stream; and the 2^22 MSM with its scalars uploaded
    from pinned host memory first (the rate a caller handing host scalars
    sees; the bases stay resident as the SRS does)."""
    import numpy as np
    import torch
    from zkprover import device as dv
    from zkprover import hostmem
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                   ctypes.c_void_p]
    nbytes = hostmem.CHUNK_BYTES
    pin = hostmem.PinnedBuffer(nbytes)
    page = np.ones(nbytes, dtype=np.uint8)
    pin.array(np.uint8)[:] = 1
    dev = torch.empty(nbytes, dtype=torch.uint8, device=dv.device())
    st = dv.stream_handle()
    res = {"bytes": nbytes}
    try:
        for name, dst, src, kind in (("h2d_pinned", dev.data_ptr(), pin.ptr, 1),
                                     ("d2h_pinned", pin.ptr, dev.data_ptr(), 2),
                                     ("h2d_pageable", dev.data_ptr(), page.ctypes.data, 1),
                                     ("d2h_pageable", page.ctypes.data, dev.data_ptr(), 2)):
            if hip.hipMemcpyAsync(dst, src, nbytes, kind, st) != 0:
                raise RuntimeError("hipMemcpyAsync " + name)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            t0 = time.perf_counter()
            e0.record()
            for _ in range(reps):
                if hip.hipMemcpyAsync(dst, src, nbytes, kind, st) != 0:
                    raise RuntimeError("hipMemcpyAsync " + name)
            e1.record()
            torch.cuda.synchronize()
            ms = max(e0.elapsed_time(e1), (time.perf_counter() - t0) * 1e3) / reps
            res[name + "_GBps"] = round(nbytes / (ms * 1e-3) / 1e9, 2)
    finally:
        pin.close()
    upload_ms = nbytes / (res["h2d_pinned_GBps"] * 1e9) * 1e3
    res["msm_2^22_with_pinned_scalar_upload_ms"] = round(msm_ms + upload_ms, 3)
    return res


class _MsmItem(ctypes.Structure):
    _fields_ = [("points", ctypes.c_void_p), ("scalars", ctypes.c_void_p), ("n", ctypes.c_int),
                ("table_c", ctypes.c_int)]


def bench_msm_table(L, reps=3, count=16):
    """Fixed-base tables against plain points for the sizes the prover core
    builds them for (its keys hold tables up to 2^20 points): per size, one
    batch of `count` MSMs (the prover core's commitment group) over shared
    bases through zk_msm_batch, plain at the default window and against the
    table at zk_msm_table_window; host clock, synchronous calls, identical
    results required.  The table is built once per size (set-up time
    reported beside it)."""
    import numpy as np
    import torch
    from zkprover import device as dv
    from zkprover._lib import call
    dev = dv.device()
    st = dv.stream_handle()
    L.zk_msm_table_window.restype = ctypes.c_int
    L.zk_msm_table_window.argtypes = [ctypes.c_int]
    L.zk_msm_table_bytes.restype = ctypes.c_size_t
    L.zk_msm_table_bytes.argtypes = [ctypes.c_int, ctypes.c_int]
    L.zk_msm_table_build.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    L.zk_msm_batch_scratch_bytes.restype = ctypes.c_size_t
    L.zk_msm_batch_scratch_bytes.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.zk_msm_batch.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                               ctypes.c_size_t, ctypes.c_void_p]
    res = {}
    for lg in (16, 18, 20):
        n = 1 << lg
        pts = torch.empty(n * 64, dtype=torch.uint8, device=dev)
        sc = torch.empty(count * n * 32, dtype=torch.uint8, device=dev)
        call("zk_gen_points", pts.data_ptr(), n, 40 + lg, st)
        call("zk_rand_fr", sc.data_ptr(), count * n, 50 + lg, st)
        tc = L.zk_msm_table_window(n)
        table = torch.empty(L.zk_msm_table_bytes(n, tc), dtype=torch.uint8, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        call("zk_msm_table_build", table.data_ptr(), pts.data_ptr(), n, tc, st)
        torch.cuda.synchronize()
        build_ms = (time.perf_counter() - t0) * 1e3
        outs, per = {}, {}
        for name, src, c in (("plain", pts, 0), ("table", table, tc)):
            items = (_MsmItem * count)(*[_MsmItem(src.data_ptr(), sc.data_ptr() + m * n * 32, n, c)
                                         for m in range(count)])
            nb = L.zk_msm_batch_scratch_bytes(ctypes.addressof(items), count, 0)
            scratch = torch.empty(nb, dtype=torch.uint8, device=dev)
            out = torch.zeros(count * 65536, dtype=torch.uint8, device=dev)

            def run():
                call("zk_msm_batch", out.data_ptr(), ctypes.addressof(items), count, 0, scratch.data_ptr(), nb, st)
            run()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                run()
            torch.cuda.synchronize()
            per[name] = (time.perf_counter() - t0) * 1e3 / reps / count
            outs[name] = out.view(count, 65536)[:, :96].cpu().numpy().copy()
            del scratch, out
        res[f"2^{lg}"] = {"plain_ms_per_msm": round(per["plain"], 4), "table_ms_per_msm": round(per["table"], 4),
                          "speedup": round(per["plain"] / per["table"], 3), "plain_window": L.zk_msm_default_window(n),
                          "table_window": tc, "table_bytes": table.numel(), "build_ms": round(build_ms, 2),
                          "identical_results": bool((outs["plain"] == outs["table"]).all())}
        del pts, sc, table
        torch.cuda.empty_cache()
    res["workload"] = f"{count} MSMs per size over shared bases, uniform scalars (one commitment group)"
    return res


def bench_batched_msm(L, points, reps):
    """64 independent MSMs with sizes log-uniform in [2^12, 2^20] over one
    shared base array: one batched pass (zk_msm_batch) against the plain path
    (zk_msm once per MSM, each a full pipeline + tail).  Both are synchronous
    host calls, timed with the host clock; results must be identical."""
    import numpy as np
    import torch
    from zkprover import device as dv
    from zkprover._lib import call
    rng = np.random.default_rng(2024)
    sizes = [int(round(2.0 ** x)) for x in rng.uniform(12, 20, 64)]
    total = sum(sizes)
    dev = dv.device()
    st = dv.stream_handle()
    scal = torch.empty(total * 32, dtype=torch.uint8, device=dev)
    call("zk_rand_fr", scal.data_ptr(), total, 21, st)
    items = (_MsmItem * 64)()
    off = 0
    for m, k in enumerate(sizes):
        items[m] = _MsmItem(points.data_ptr(), scal.data_ptr() + off * 32, k, 0)
        off += k
    L.zk_msm_batch_scratch_bytes.restype = ctypes.c_size_t
    L.zk_msm_batch_scratch_bytes.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.zk_msm_scratch_bytes.restype = ctypes.c_size_t
    L.zk_msm_scratch_bytes.argtypes = [ctypes.c_int, ctypes.c_int]
    L.zk_msm_batch.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                               ctypes.c_size_t, ctypes.c_void_p]
    L.zk_msm.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                                 ctypes.c_void_p]
    bb = L.zk_msm_batch_scratch_bytes(ctypes.addressof(items), 64, 0)
    pb = max(L.zk_msm_scratch_bytes(k, 0) for k in sizes)
    scratch = torch.empty(max(bb, pb), dtype=torch.uint8, device=dev)
    out_b = torch.zeros(64 * 65536, dtype=torch.uint8, device=dev)
    out_p = torch.zeros(64 * 65536, dtype=torch.uint8, device=dev)

    def batched():
        call("zk_msm_batch", out_b.data_ptr(), ctypes.addressof(items), 64, 0, scratch.data_ptr(), bb, st)

    def plain():
        for m in range(64):
            call("zk_msm", out_p.data_ptr() + m * 65536, items[m].points, items[m].scalars, items[m].n, 0,
                 scratch.data_ptr(), pb, st)

    res = {}
    for name, fn in (("plain", plain), ("batched", batched)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0) * 1e3 / reps
    same = bool(torch.equal(out_b.view(64, 65536)[:, :96], out_p.view(64, 65536)[:, :96]))
    return {"workload": "64 G1 MSMs, sizes log-uniform 2^12..2^20 (%d points), shared bases" % total,
            "plain_ms": round(res["plain"], 3), "batched_ms": round(res["batched"], 3),
            "speedup": round(res["plain"] / res["batched"], 3), "identical_results": same}


def bench_proof_core(ks, reps=5):
    """One proof's device arithmetic (prover core) for the benchmark circuit
    family (40 advice, 10 fixed, 1 instance column, 6 range lookups into a
    2^16 table, 12 permutation columns, 2 shuffles of up to 2^18 rows) at
    each k, after one warm-up proof: the median of `reps` proofs through the
    batched MSM, through one MSM call per polynomial (both opened with GWC),
    batched and opened with Shplonk, batched with the h(X) round replayed
    from a HIP graph, batched with the advice and instance columns uploaded
    from pinned host memory during the commitment phase (as
    create_proof_from_advices receives them), and batched after the MSM
    window and NTT radix autotuners ran for the proof's sizes.  The device-witness modes
    exclude the witness synthesis (the reference takes the advice as input)."""
    import statistics

    import torch
    from zkprover import hostmem
    from zkprover import autotune
    from zkprover.prover import BATCHED, GRAPH, HOST_WITNESS, SHPLONK, Circuit, Prover, ProvingKey, footprint
    res = {}
    for k in ks:
        progress("proof core k=%d" % k)
        circ = Circuit.benchmark(k)
        pk_b, job_b = footprint(circ)
        pk = ProvingKey(circ, seed=7)
        pr = Prover(pk)
        bufs = []
        try:
            pr.prove(1, flags=BATCHED, want_pieces=False)
            out, ok = {}, True
            pr.prove(1, flags=BATCHED | GRAPH, want_pieces=False)  # captures the h(X) round
            for mode, flags in (("batched", BATCHED), ("plain", 0), ("shplonk_batched", BATCHED | SHPLONK),
                                ("graph_batched", BATCHED | GRAPH)):
                t, info = [], None
                for r in range(reps):
                    _, _, info = pr.prove(2 + r, flags=flags, want_pieces=False)
                    t.append(info.total_ms)
                    ok = ok and bool(info.z_ok and info.quotient_ok and info.open_ok)
                out[mode] = (statistics.median(t), info)
            n = 1 << k
            bufs = [hostmem.PinnedBuffer(n * 32) for _ in range(circ.advice + circ.instance)]
            adv = [b.ptr for b in bufs[:circ.advice]]
            ins = [b.ptr for b in bufs[circ.advice:]]
            pr.prove(2, flags=BATCHED, want_pieces=False)
            pr.witness_to_host(adv, ins)
            pr.set_host_witness(adv, ins)
            t = []
            for r in range(reps):
                _, _, info = pr.prove(2, flags=BATCHED | HOST_WITNESS, want_pieces=False)
                t.append(info.total_ms)
                ok = ok and bool(info.z_ok and info.quotient_ok and info.open_ok)
            out["host_witness"] = (statistics.median(t), info)
            # the autotuners for this proof's MSM and NTT sizes, then the
            # batched proof again; the default table is restored afterwards
            progress("autotune k=%d" % k)
            tuned_msm = autotune.tune_msm([k], reps=2)
            tuned_ntt = autotune.tune_ntt([k, k + EXT], reps=3)
            t = []
            for r in range(reps):
                _, _, info = pr.prove(2 + r, flags=BATCHED, want_pieces=False)
                t.append(info.total_ms)
                ok = ok and bool(info.z_ok and info.quotient_ok and info.open_ok)
            out["tuned_batched"] = (statistics.median(t), info)
            out["tuned_choice"] = {"msm_window": {kk: v[0] for kk, v in tuned_msm.items()},
                                   "ntt_max_deg": {kk: v[0] for kk, v in tuned_ntt.items()}}
            from zkprover._lib import call as _call
            _call("zk_msm_set_window", k, 0)
            for kk in (k, k + EXT):
                _call("zk_ntt_set_max_deg", kk, 0)
        finally:
            for b in bufs:
                b.close()
            pr.close()
            pk.close()
            torch.cuda.synchronize()
        ib = out["batched"][1]
        res["k%d" % k] = {"batched_s": round(out["batched"][0] / 1e3, 4), "plain_s": round(out["plain"][0] / 1e3, 4),
                          "speedup": round(out["plain"][0] / out["batched"][0], 3),
                          "shplonk_batched_s": round(out["shplonk_batched"][0] / 1e3, 4),
                          "host_witness_batched_s": round(out["host_witness"][0] / 1e3, 4),
                          "graph_batched_s": round(out["graph_batched"][0] / 1e3, 4),
                          "tuned_batched_s": round(out["tuned_batched"][0] / 1e3, 4),
                          "tuned_choice": out["tuned_choice"],
                          "h_eval_ms": {"eager": round(ib.phase_ms[4], 3),
                                        "graph": round(out["graph_batched"][1].phase_ms[4], 3)},
                          "proofs_per_mode": reps, "statistic": "median",
                          "phases_ms": ib.as_dict()["phase_ms"],
                          "host_witness_phases_ms": out["host_witness"][1].as_dict()["phase_ms"],
                          "shplonk_open_ms": round(out["shplonk_batched"][1].phase_ms[6], 3),
                          "all_checks_ok": ok,
                          "device_bytes": {"key": pk_b, "per_proof": job_b},
                          "host_witness_bytes": (circ.advice + circ.instance) * (1 << k) * 32}
    return res


def bench_serving(count):
    """Proofs of k in {18, 20, 22} (weights 0.5 / 0.3 / 0.2) arriving as a
    Poisson stream, offered at 0.5, 1, 2 and 4x the one-at-a-time capacity:
    the plain path (one slot, FIFO: one proof at a time, each with its own
    batched MSMs) against the server (3 slots, per-proof states pooled at
    creation, a proof joining those in flight when more than two requests
    wait per running proof), same arrivals and seeds; and the server with the
    commitment MSMs of concurrent proofs merged into shared batched passes
    (ZKS_CROSS_BATCH) at the two saturated loads.  For each plain operating
    point, the served throughput reachable at the same or a lower p99
    latency is reported beside it."""
    from zkprover.prover import BATCHED, Circuit
    from zkprover.server import CROSS_BATCH, Server, run_poisson, summarize
    shapes = [Circuit.benchmark(k) for k in (18, 20, 22)]
    weights = [0.5, 0.3, 0.2]
    all_loads = [0.5, 1.0, 2.0, 4.0]
    out = {"workload": "Poisson arrivals of k=18/20/22 proofs of the benchmark circuit (weights 0.5/0.3/0.2), "
                       "%d requests per load" % count}
    res = {}
    for mode, slots, flags, loads in (("plain", 1, BATCHED, all_loads), ("served", 3, BATCHED, all_loads),
                                      ("served_cross_batch", 3, BATCHED | CROSS_BATCH, [2.0, 4.0])):
        srv = Server(shapes, slots=slots, flags=flags)
        try:
            t = []
            for sh in range(3):
                srv.wait(srv.submit(sh, 1), sh)
                _, r = srv.wait(srv.submit(sh, 2), sh)
                t.append(r.run_ms)
            if mode == "plain":
                mean_ms = sum(w * x for w, x in zip(weights, t))
                out["single_proof_ms"] = {"k18": round(t[0], 1), "k20": round(t[1], 1), "k22": round(t[2], 1)}
                out["capacity_per_s"] = round(1000.0 / mean_ms, 3)
            res[mode] = {}
            for ld in loads:
                progress("serving %s load %.1f" % (mode, ld))
                recs, wall = run_poisson(srv, ld * 1000.0 / mean_ms, count, weights, seed=10)
                res[mode]["load_%.1f" % ld] = summarize(recs, wall)
            res[mode]["server"] = srv.stats()
            res[mode]["batcher"] = srv.batch_stats()
        finally:
            srv.close()
    out.update(res)
    # for every plain operating point: the served throughput reachable with a
    # p99 no worse than the plain path's there
    cmp = []
    for k, v in res["plain"].items():
        if not k.startswith("load_"):
            continue
        ok = [w["proofs_per_hour"] for kk, w in res["served"].items()
              if kk.startswith("load_") and w["p99_ms"] <= v["p99_ms"]]
        cmp.append({"load": k, "p99_ms": v["p99_ms"], "plain_proofs_per_hour": v["proofs_per_hour"],
                    "served_proofs_per_hour": max(ok) if ok else None})
    out["at_equal_p99"] = cmp
    return out


class _NttItem(ctypes.Structure):
    _fields_ = [("buf", ctypes.c_void_p), ("tmp", ctypes.c_void_p), ("omegas", ctypes.c_void_p),
                ("n_inv", ctypes.c_void_p), ("log_n", ctypes.c_int), ("reserved", ctypes.c_int)]


def bench_batched_ntt(L, reps):
    """64 independent forward NTTs with log-sizes uniform in [14, 21]: one
    zk_ntt_batch call (one launch per pass round) against 64 `ntt` calls.
    Both are asynchronous; timed with HIP events on the launch stream."""
    import numpy as np
    import torch
    from zkprover import device as dv
    from zkprover._lib import call
    rng = np.random.default_rng(2025)
    logs = [int(x) for x in rng.integers(14, 22, 64)]
    dev = dv.device()
    st = dv.stream_handle()
    tables = {}
    for k in sorted(set(logs)):
        init = np.zeros((1 << k, 4), dtype=np.uint64)
        init[0] = _to_mont_limbs(1)
        init[1] = _to_mont_limbs(_root_of_unity_int(k))
        t = dv.to_dev(init)
        call("expand_omega_buffer", t.data_ptr(), k, 1)
        tables[k] = t
    bufs = []
    items = (_NttItem * 64)()
    for m, k in enumerate(logs):
        b = torch.empty((1 << k) * 32, dtype=torch.uint8, device=dev)
        t = torch.empty_like(b)
        call("zk_rand_fr", b.data_ptr(), 1 << k, 700 + m, st)
        bufs.append((b, t))
        items[m] = _NttItem(b.data_ptr(), t.data_ptr(), tables[k].data_ptr(), None, k, 0)
    swapped = (ctypes.c_ubyte * 64)()
    swap = ctypes.c_bool(False)
    L.zk_ntt_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]

    def batched():
        call("zk_ntt_batch", ctypes.addressof(items), 64, ctypes.addressof(swapped), st)

    def plain():
        for m in range(64):
            call("ntt", items[m].buf, items[m].tmp, None, items[m].omegas, None, items[m].log_n, 8,
                 ctypes.addressof(swap), st)

    res = {}
    for name, fn in (("plain", plain), ("batched", batched)):
        fn()
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            fn()
        e.record()
        torch.cuda.synchronize()
        res[name] = s.elapsed_time(e) / reps
    # identical transforms: both paths once more from the same inputs (the
    # result buffer of a transform depends only on its pass count)
    outs = []
    for fn in (plain, batched):
        for m, (b, t) in enumerate(bufs[:4]):
            call("zk_rand_fr", b.data_ptr(), 1 << logs[m], 900 + m, st)
        fn()
        torch.cuda.synchronize()
        outs.append([(t if swapped[m] else b).clone() for m, (b, t) in enumerate(bufs[:4])])
    same = all(torch.equal(a, b) for a, b in zip(*outs))
    elems = sum(1 << k for k in logs)
    return {"workload": "64 Fr NTTs, log-sizes uniform in [14, 21] (%d elements)" % elems,
            "plain_ms": round(res["plain"], 3), "batched_ms": round(res["batched"], 3),
            "speedup": round(res["plain"] / res["batched"], 3), "identical_results": same}


_R = 21888242871839275222246405745257275088548364400416034343698204186575808495617


def _root_of_unity_int(log_n):
    w = pow(7, (_R - 1) >> 28, _R)
    for _ in range(log_n, 28):
        w = w * w % _R
    return w


def _to_mont_limbs(v):
    m = (v << 256) % _R
    return [(m >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(4)]


def _pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by what the counter reads
    for a plain 256 MiB copy); None when no summary is present"""
    path = os.path.join(ROOT, "profiles", "msm_2e22_pmc_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 2 and parts[0] == kernel:
            try:
                return int(float(parts[-1]))
            except ValueError:
                return None
    return None


def _cpu_info():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"nproc": os.cpu_count(), "usable_cpus": usable, "cpu_model": model}


def cpu_baseline(n):
    """The C oracle on the host cores this job is given (OMP_NUM_THREADS, else
    every usable CPU), one full run of each primitive at the zkWasm shape:
    G1 MSM n = 2^22 (Pippenger parallel over point ranges x windows), Fr NTT
    2^22, coset NTT 2^22 -> 2^24, batch inversion 2^22 (parallel chunks); and
    a lower bound for one k=20 benchmark proof's arithmetic: the proof's
    primitive counts times the oracle's time for each primitive at k=20 (it
    leaves out the h(X) evaluation, the grand products' element-wise passes
    and the openings' scans)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import zkoracle as zo
    info = _cpu_info()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or info["usable_cpus"] or 1
    zo.set_threads(threads)

    def timed(fn):
        t0 = time.perf_counter()
        fn()
        return (time.perf_counter() - t0) * 1e3

    def prims(log_n):
        m = 1 << log_n
        pts = zo.gen_bases_ap(m, 4)
        sc = zo.rand_fr(m, 3)
        x = zo.rand_fr(m, 5)
        w = zo.root_of_unity(log_n)
        we = zo.root_of_unity(log_n + EXT)
        zeta = pow(7, (_R - 1) // 3, _R)
        cp = zo.to_mont([zeta, zeta * zeta % _R])
        ext = zo.fr_zeros(m << EXT)
        inv_in = zo.rand_fr_nonzero(m, 7)
        L = zo.lib()
        xa = np.array(x, copy=True)
        out = {"msm": timed(lambda: zo.msm(pts, sc)),
               "ntt": timed(lambda: L.zo_ntt(zo._p(xa), ctypes.c_size_t(log_n), zo._p(w), None))}
        ext[:m] = x

        def coset():
            # in place: scale by the coset powers, then the 2^(log_n + EXT) NTT
            L.zo_extended_prepare(zo._p(ext), zo._p(cp), ctypes.c_int(3), ctypes.c_size_t(m),
                                  ctypes.c_size_t(m << EXT), ctypes.c_int(1))
            L.zo_ntt(zo._p(ext), ctypes.c_size_t(log_n + EXT), zo._p(we), None)
        out["coset"] = timed(coset)
        out["batch_inversion"] = timed(lambda: L.zo_batch_invert(zo._p(inv_in), ctypes.c_size_t(m)))
        return out

    p22 = prims(LOG_N)
    p20 = prims(20)
    from zkprover.prover import Circuit
    c = Circuit.benchmark(20)
    Sp = (c.perm_cols + 1) // 2
    counts = {"msm": c.instance + c.advice + 2 * c.lookups + Sp + c.shuffles + 1 + 3 + 3,
              "ntt": c.instance + c.advice + 2 * c.lookups + Sp + c.shuffles + 1,
              "coset": 4 + 3 * c.lookups + c.perm_cols + Sp + 3 * c.shuffles + (c.fixed - 2),
              "batch_inversion": 2 * c.lookups + Sp + c.shuffles}
    est = sum(counts[k] * p20[k] for k in counts)
    return {"value": round(p22["msm"], 1), "unit": "ms (G1 MSM, n=2^22)", "cores": threads, "kind": "port",
            "cpu": info,
            "sample": "one full run of each primitive at n=2^22 on the C oracle (the MSM is `value`)",
            "primitives_ms": {"msm_2^22": round(p22["msm"], 1), "ntt_2^22": round(p22["ntt"], 1),
                              "coset_ntt_2^22_to_2^24": round(p22["coset"], 1),
                              "batch_inversion_2^22": round(p22["batch_inversion"], 1)},
            "primitives_k20_ms": {k: round(v, 1) for k, v in p20.items()},
            "proof_k20_lower_bound_s": round(est / 1e3, 2),
            "proof_k20_counts": counts}


if __name__ == "__main__":
    main()
