"""Loads libzkwasm_prover_kernel.so and declares the C-ABI signatures.

There is no fallback: if the HIP library is missing or fails to load, every
call raises.  The declared symbols mirror include/zkwasm_prover_kernel.h.
"""
import ctypes
import os

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_DIR, "libzkwasm_prover_kernel.so")

_lib = None

P = ctypes.c_void_p
I32 = ctypes.c_int
U32 = ctypes.c_uint
SZ = ctypes.c_size_t

# name -> argtypes (all return hipError_t as int)
SIGNATURES = {
    # drop-in symbols (src/cuda/bn254_c.rs)
    "ntt": [P, P, P, P, P, I32, I32, P, P],
    "expand_omega_buffer": [P, I32, I32],
    "extended_prepare": [P, P, U32, I32, I32, I32, P],
    "field_op": [P, P, I32, P, P, I32, P, I32, I32, P],
    "field_op_batch_mul_sum": [P, P, P, I32, I32, P],
    "field_mul_zip": [P, P, I32, I32],
    "field_sum": [P, P, P, P, P, I32, I32],
    "lookup_eval_h": [P, P, P, P, P, P, P, P, P, P, P, P, I32, I32, P],
    "permutation_eval_h_p1": [P, P, P, P, P, P, I32],
    "permutation_eval_h_p2": [P, P, P, P, P, I32, I32, I32],
    "permutation_eval_h_l": [P, P, P, P, I32, P],
    "logup_eval_h": [P, P, P, P, P, P, P, P, P, P, P, I32, I32, P],
    "logup_eval_h_extra_inputs": [P, P, P, P, P, P, I32, I32, P],
    "logup_eval_h_z_set": [P, P, P, P, P, I32, I32, I32, P],
    "logup_eval_h_inputs_product_sum": [P, P, P, I32, I32, P],
    "logup_sum_input_inv": [P, P, P, P, I32, I32, P],
    "shuffle_eval_h": [P, P, P, P, P, P, P, P, I32, I32],
    "shuffle_eval_h_v2": [P, P, P, P, I32, P, P, P, P, P, I32, I32],
    "eval_logup_z_compose": [P, P, P, P, P, P, I32, I32, P],
    "eval_logup_z_pure": [P, P, P, P, I32, I32, P],
    "prepare_table_lookup": [P, P, P, P, P, P, P, U32, U32, U32, P],
    "calc_m": [P, P, P, P, P, P, P, P, U32, U32, U32, P],
    "poly_eval": [P, P, P, I32, P],
    "shplonk_h_x_merge": [P, P, P, P, P, I32, I32],
    "shplonk_h_x_div_points": [P, P, P, I32, I32],
    "msm": [P, P, P, P, P, P, P, P, P, I32, I32, I32, I32, I32, I32, I32, P],
    "batch_msm_collect": [P, P, P, P, P, U32, U32, U32, U32, P],
    # extensions (no reference counterpart)
    "zk_batch_invert": [P, P, I32, P],
    "zk_kate_divide": [P, P, P, P, I32, P, P],
    "zk_shplonk_h_x_merge_on": [P, P, P, P, P, I32, I32, P],
    "zk_shplonk_h_x_div_points_on": [P, P, P, I32, I32, P],
    "zk_permutation_z_set": [P, P, P, I32, P, P, P, P, P, P, I32, P],
    "zk_permutation_chain": [P, I32, I32, I32, P, P],
    "zk_shuffle_z": [P, P, P, I32, P, P, I32, P],
    "zk_permutation_eval_h_set": [P, P, P, P, I32, P, P, P, P, P, P, I32, I32, P],
    # element-wise / calibration entry points used by tests and bench
    "zk_fr_op": [P, P, P, I32, I32, P],
    "zk_fp_op": [P, P, P, I32, I32, P],
    "zk_g1_op": [P, P, P, I32, I32, P],
    "zk_fp29_op": [P, P, P, I32, I32, P],
    "zk_bench_mul29": [P, I32, I32, I32, I32, P],
    "zk_bench_mul": [P, I32, I32, I32, P],
    "zk_bench_dbl": [P, I32, I32, I32, P],
    "zk_bench_madd": [P, P, I32, I32, I32, I32, P],
    "zk_copy": [P, P, SZ, P],
    "zk_rand_fr": [P, I32, ctypes.c_ulonglong, P],
    "zk_gen_points": [P, I32, ctypes.c_ulonglong, P],
    # MSM / NTT with caller scratch, batched forms, tuning tables
    "zk_msm": [P, P, P, I32, I32, P, SZ, P],
    "zk_msm_batch": [P, P, I32, I32, P, SZ, P],
    "zk_msm_batch_async": [P, P, I32, I32, P, SZ, P],
    "zk_graph_begin": [P],
    "zk_graph_end": [P, P],
    "zk_graph_launch": [P, P],
    "zk_graph_destroy": [P],
    "zk_ntt_batch": [P, I32, P, P],
    "zk_msm_set_window": [I32, I32],
    "zk_msm_autotune": [P, P, I32, I32, P, P, P],
    "zk_ntt_set_max_deg": [I32, I32],
    "zk_ntt_autotune": [P, P, P, I32, I32, P, P, P],
}


class HipError(RuntimeError):
    pass


def lib():
    """The loaded kernel library; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HipError(f"HIP kernel library not built: {LIB_PATH} (run zkwasm-prover_amd/build.py)")
        _lib = ctypes.CDLL(LIB_PATH)
        for name, argtypes in SIGNATURES.items():
            fn = getattr(_lib, name, None)
            if fn is None:
                continue
            fn.argtypes = argtypes
            fn.restype = ctypes.c_int
    return _lib


def exported_symbols():
    l = lib()
    return [name for name in SIGNATURES if hasattr(l, name)]


def check(err, what):
    if err != 0:
        raise HipError(f"{what} failed with hipError {err}")


def call(name, *args):
    fn = getattr(lib(), name)
    check(fn(*args), name)
