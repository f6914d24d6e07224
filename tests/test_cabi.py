"""CPU: the C-ABI library loads and exports every symbol include/*.h declares,
and the host-side logic mirrors the reference's (no GPU calls)."""
import ctypes
import math
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER_DIR = os.path.join(ROOT, "include")

# symbols of src/cuda/bn254_c.rs:5-309 that are on the hot path
REFERENCE_HOT_PATH = [
    "ntt", "expand_omega_buffer", "extended_prepare", "msm", "batch_msm_collect", "logup_sum_input_inv",
    "eval_logup_z_compose", "eval_logup_z_pure", "field_op_batch_mul_sum", "field_op", "field_mul_zip",
    "permutation_eval_h_p1", "permutation_eval_h_p2", "permutation_eval_h_l", "logup_eval_h",
    "logup_eval_h_extra_inputs", "logup_eval_h_z_set", "logup_eval_h_inputs_product_sum", "shuffle_eval_h",
    "shuffle_eval_h_v2",
]


def declared_symbols():
    names = []
    for f in sorted(os.listdir(HEADER_DIR)):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(HEADER_DIR, f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:hipError_t|void|size_t|int)\s+(\w+)\s*\(", text, flags=re.M)
    return names


@pytest.fixture(scope="module")
def lib():
    from zkprover._lib import LIB_PATH
    if not os.path.exists(LIB_PATH):
        pytest.fail(f"kernel library not built: {LIB_PATH}")
    return ctypes.CDLL(LIB_PATH)


def test_header_declares_reference_hot_path():
    decl = set(declared_symbols())
    missing = [s for s in REFERENCE_HOT_PATH if s not in decl]
    assert not missing, missing


def test_library_exports_every_declared_symbol(lib):
    decl = declared_symbols()
    assert len(decl) >= len(REFERENCE_HOT_PATH)
    missing = [s for s in decl if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_table_matches_header():
    from zkprover._lib import SIGNATURES
    decl = set(declared_symbols())
    assert set(SIGNATURES) <= decl, sorted(set(SIGNATURES) - decl)


def _ref_policy(count, length):
    """src/cuda/msm.rs:274-283 (ark_std::log2 is a ceiling)"""
    def log2c(x):
        return math.ceil(math.log2(x)) if x > 1 else 0
    wb = min(log2c(count), 3) + 10 + max(log2c(length), 22) - 22
    windows = (254 + wb - 1) // wb
    workers = (length + 63) // 64 * 64 if length < 128 * 512 else 128 * 512
    return wb, windows, workers


@pytest.mark.parametrize("count,length", [(1, 1 << 16), (1, 1 << 22), (2, 1 << 22), (3, 1000), (8, 1 << 20),
                                          (64, 1 << 12), (5, (1 << 23) + 1), (1, 1 << 24)])
def test_msm_window_policy_matches_reference(lib, count, length):
    wb, w, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib.zkp_msm_policy.restype = None
    lib.zkp_msm_policy(count, length, ctypes.byref(wb), ctypes.byref(w), ctypes.byref(k))
    assert (wb.value, w.value, k.value) == _ref_policy(count, length)


def test_msm_scratch_and_window_queries(lib):
    lib.zk_msm_scratch_bytes.restype = ctypes.c_size_t
    lib.zk_msm_scratch_bytes.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.zk_msm_default_window.restype = ctypes.c_int
    sizes = [lib.zk_msm_scratch_bytes(1 << k, 0) for k in range(4, 23)]
    assert all(b > a for a, b in zip(sizes, sizes[1:]))
    assert lib.zk_msm_scratch_bytes(0, 0) == 0
    assert lib.zk_msm_default_window(1 << 22) == 17
    assert all(2 <= lib.zk_msm_default_window(1 << k) <= 17 for k in range(0, 26))


class _Item(ctypes.Structure):
    _fields_ = [("points", ctypes.c_void_p), ("scalars", ctypes.c_void_p), ("n", ctypes.c_int),
                ("table_c", ctypes.c_int)]


def test_msm_batch_scratch_query(lib):
    """batch sizing: one item sizes like zk_msm; more items need more; a
    non-positive size or count is rejected with 0"""
    f = lib.zk_msm_batch_scratch_bytes
    f.restype = ctypes.c_size_t
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    lib.zk_msm_scratch_bytes.restype = ctypes.c_size_t
    lib.zk_msm_scratch_bytes.argtypes = [ctypes.c_int, ctypes.c_int]
    for n in (1, 1000, 1 << 16, 1 << 20):
        one = (_Item * 1)(_Item(None, None, n, 0))
        for c in (0, 8, 13):
            assert f(ctypes.addressof(one), 1, c) == lib.zk_msm_scratch_bytes(n, c)
    two = (_Item * 2)(_Item(None, None, 1 << 16, 0), _Item(None, None, 1 << 16, 0))
    one = (_Item * 1)(_Item(None, None, 1 << 16, 0))
    assert f(ctypes.addressof(two), 2, 12) > f(ctypes.addressof(one), 1, 12)
    bad = (_Item * 2)(_Item(None, None, 1 << 16, 0), _Item(None, None, 0, 0))
    assert f(ctypes.addressof(bad), 2, 0) == 0
    assert f(ctypes.addressof(one), 0, 0) == 0
    lib.zk_msm_batch.restype = ctypes.c_int
    lib.zk_msm_batch.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                 ctypes.c_size_t, ctypes.c_void_p]
    assert lib.zk_msm_batch(None, ctypes.addressof(one), 1, 0, None, 0, None) != 0
    assert lib.zk_msm_batch(ctypes.addressof(one), ctypes.addressof(bad), 2, 0, ctypes.addressof(one), 16,
                            None) != 0


def test_msm_table_sizing_host_only(lib):
    """fixed-base tables: ceil(255 / c) rows of 64 B per point for 8 <= c <= 20;
    the window model grows with n up to c = 17 (wider windows sort slower than
    the additions they save); a batch
    mixing plain and table items needs the larger pass's scratch, and a table
    item with an invalid window has none"""
    lib.zk_msm_table_bytes.restype = ctypes.c_size_t
    lib.zk_msm_table_bytes.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.zk_msm_table_window.restype = ctypes.c_int
    lib.zk_msm_table_window.argtypes = [ctypes.c_int]
    for c in range(8, 21):
        assert lib.zk_msm_table_bytes(1000, c) == -(-255 // c) * 1000 * 64
    for c in (0, 7, 21, 32):
        assert lib.zk_msm_table_bytes(1000, c) == 0
    assert lib.zk_msm_table_bytes(0, 12) == 0 and lib.zk_msm_table_bytes((1 << 31) - 1, 8) == 0
    assert lib.zk_msm_table_window(1 << 16) == 16 and lib.zk_msm_table_window(1 << 22) == 17
    ws = [lib.zk_msm_table_window(1 << k) for k in range(4, 27)]
    assert all(8 <= w <= 17 for w in ws) and ws == sorted(ws)
    f = lib.zk_msm_batch_scratch_bytes
    f.restype = ctypes.c_size_t
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    n = 1 << 16
    plain = (_Item * 1)(_Item(None, None, n, 0))
    tab = (_Item * 1)(_Item(None, None, n, 16))
    mixed = (_Item * 2)(_Item(None, None, n, 0), _Item(None, None, n, 16))
    p, t, m = (f(ctypes.addressof(x), len(x), 0) for x in (plain, tab, mixed))
    assert p > 0 and t > 0 and m == max(p, t)
    bad = (_Item * 1)(_Item(None, None, n, 21))
    assert f(ctypes.addressof(bad), 1, 0) == 0


def test_invalid_arguments_rejected_without_gpu(lib):
    """argument checks happen on the host before any launch"""
    lib.ntt.restype = ctypes.c_int
    assert lib.ntt(None, None, None, None, None, -1, 8, None, None) != 0
    lib.field_op.restype = ctypes.c_int
    assert lib.field_op(None, None, 0, None, None, 0, None, 64, 1, None) != 0
    lib.zk_msm.restype = ctypes.c_int
    lib.zk_msm.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                                     ctypes.c_void_p]
    assert lib.zk_msm(None, None, None, 0, 0, None, 0, None) != 0
    lib.eval_logup_z_pure.restype = ctypes.c_int
    assert lib.eval_logup_z_pure(None, None, None, None, 5, 4, None) != 0


def test_helper_entry_points_reject_bad_arguments_without_gpu(lib):
    """the element-wise helpers and calibration launches check buffers, op
    codes and launch shapes on the host (a zero block would otherwise divide
    by zero, a negative iteration count would run for 2^32 steps); the
    non-null dummy address is never dereferenced on these paths"""
    v, i = ctypes.c_void_p, ctypes.c_int
    dummy = ctypes.c_void_p(ctypes.addressof(ctypes.create_string_buffer(64)))
    for f, args in (("zk_fr_op", [None, dummy, None, 1, 4, None]), ("zk_fr_op", [dummy, dummy, None, 8, 4, None]),
                    ("zk_fp_op", [dummy, None, None, 0, 4, None]), ("zk_fp_op", [dummy, dummy, None, -1, 4, None]),
                    ("zk_fp29_op", [dummy, dummy, None, 5, 4, None]), ("zk_g1_op", [dummy, dummy, None, 0, 4, None]),
                    ("zk_g1_op", [dummy, dummy, dummy, 5, 4, None]),
                    ("zk_bench_mul", [dummy, 256, 0, 4, None]), ("zk_bench_mul", [dummy, 256, 96, 4, None]),
                    ("zk_bench_mul", [dummy, 256, 2048, 4, None]), ("zk_bench_mul", [dummy, 256, 64, -1, None]),
                    ("zk_bench_mul", [None, 256, 64, 4, None]), ("zk_bench_mul_cios", [dummy, 0, 64, 4, None]),
                    ("zk_bench_dbl", [dummy, -64, 64, 4, None]), ("zk_bench_mul29", [dummy, 256, 0, 4, 0, None]),
                    ("zk_bench_madd", [dummy, None, 4, 256, 64, 4, None]),
                    ("zk_bench_madd", [dummy, dummy, 4, 256, 0, 4, None])):
        fn = getattr(lib, f)
        fn.restype = i
        fn.argtypes = [i if isinstance(a, int) else v for a in args]
        assert fn(*args) != 0, (f, args)
    lib.zk_copy.restype = i
    lib.zk_copy.argtypes = [v, v, ctypes.c_size_t, v]
    assert lib.zk_copy(None, dummy, 64, None) != 0
    assert lib.zk_copy(dummy, dummy, 60, None) != 0
    assert lib.zk_copy(None, None, 0, None) == 0  # nothing to copy, nothing launched


def test_tuning_tables_host_only(lib, tmp_path):
    """the per-size window / radix tables are host state: set, read back,
    reset, reject out-of-range values; autotune.save/load round-trips them"""
    lib.zk_msm_set_window.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.zk_msm_get_window.restype = ctypes.c_int
    lib.zk_ntt_set_max_deg.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.zk_ntt_get_max_deg.restype = ctypes.c_int
    lib.zk_msm_default_window.restype = ctypes.c_int
    default12 = lib.zk_msm_get_window(12)
    assert default12 == lib.zk_msm_default_window(1 << 12)
    assert lib.zk_msm_set_window(12, 9) == 0 and lib.zk_msm_get_window(12) == 9
    assert lib.zk_msm_default_window(4000) == 9  # ceil(log2 4000) = 12
    assert lib.zk_msm_set_window(12, 0) == 0 and lib.zk_msm_get_window(12) == default12
    assert lib.zk_msm_set_window(12, 18) != 0 and lib.zk_msm_set_window(40, 8) != 0
    assert lib.zk_ntt_get_max_deg(20) == 8
    assert lib.zk_ntt_set_max_deg(20, 6) == 0 and lib.zk_ntt_get_max_deg(20) == 6
    assert lib.zk_ntt_set_max_deg(20, 0) == 0 and lib.zk_ntt_get_max_deg(20) == 8
    assert lib.zk_ntt_set_max_deg(20, 9) != 0
    from zkprover import autotune
    path = str(tmp_path / "tune.json")
    autotune.save(path, msm={14: (10, {}), 16: (12, {})}, ntt={18: (7, {})})
    autotune.load(path)
    try:
        assert lib.zk_msm_get_window(14) == 10 and lib.zk_msm_get_window(16) == 12
        assert lib.zk_ntt_get_max_deg(18) == 7
        t = autotune.table()
        assert t["msm_window"][14] == 10 and t["ntt_max_deg"][18] == 7
    finally:
        lib.zk_msm_set_window(14, 0)
        lib.zk_msm_set_window(16, 0)
        lib.zk_ntt_set_max_deg(18, 0)


@pytest.mark.parametrize("count", [1, 2, 3, 8])
def test_msm_fits_reference_buffers(lib, count):
    """for every size the reference driver may pass, msm() finds a window
    whose scratch fits the index / sort buffers and the curve slice it gets
    (src/cuda/msm.rs:274-311 sizes); at 2^22 it keeps the default window,
    elsewhere at most one below it (2^18 points in a batch of 2: the
    reference's index buffers hold the buckets of c = 14, not of c = 15)"""
    wb, w, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib.zkp_msm_policy.restype = None
    lib.zk_msm_plan_window.restype = ctypes.c_int
    lib.zk_msm_default_window.restype = ctypes.c_int
    for n in [1, 2, 3, 63, 64, 65, 100, 255, 1000, 4095, 4096, 4097, 1 << 12, 1 << 14, 12345, 1 << 16,
              1 << 18, (1 << 20) + 7, 1 << 22, (1 << 23) + 1, 1 << 24]:
        lib.zkp_msm_policy(count, n, ctypes.byref(wb), ctypes.byref(w), ctypes.byref(k))
        c = lib.zk_msm_plan_window(n, w.value, wb.value, k.value)
        assert 2 <= c <= 17, (n, count)
        if n >= 1 << 12:
            assert c >= lib.zk_msm_default_window(n) - 1, (n, count, c)
        if n == 1 << 22:
            assert c == lib.zk_msm_default_window(n), (n, count, c)


def test_prover_core_sizing_host_only(lib):
    """footprints and commitment counts come from a host-side layout pass;
    inconsistent circuits are rejected with 0"""
    from zkprover.prover import Circuit, footprint
    lib.zkp_commitment_count.restype = ctypes.c_int
    lib.zkp_commitment_count.argtypes = [ctypes.c_void_p]
    c20 = Circuit.benchmark(20)
    pk20, job20 = footprint(c20)
    pk18, job18 = footprint(Circuit.benchmark(18))
    assert pk20 > pk18 > 0 and job20 > job18 > 0
    # 1 instance + 40 advice + 6 m + 6 logup z + 6 permutation sets + 2 shuffle
    # z + 1 random + 3 pieces + 3 GWC quotients (rotations last, cur, next)
    assert lib.zkp_commitment_count(ctypes.byref(c20)) == 1 + 40 + 12 + 6 + 2 + 1 + 3 + 3
    # per proof at k=20: ~ 40 * 3 n + 63 coset polys of 4n, well under 288 GB
    assert 8e9 < job20 < 64e9
    bad = Circuit.benchmark(20)
    bad.advice = 10  # fewer than the columns the roles need
    assert footprint(bad) == (0, 0)
    # every field is honoured or the shape is rejected: no table or selector
    # column, more instance than permutation columns, a negative count, or
    # coefficient gates without advice columns for their outputs
    for field, value in (("fixed", 1), ("instance", 13), ("instance", -1), ("fixed", 30)):
        bad = Circuit.benchmark(20)
        setattr(bad, field, value)
        assert footprint(bad) == (0, 0), (field, value)
        assert lib.zkp_commitment_count(ctypes.byref(bad)) == 0
