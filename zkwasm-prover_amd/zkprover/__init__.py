"""zkprover: host-side mirror of zkWasm-prover's CUDA wrappers over the HIP C-ABI.

The kernels live in libzkwasm_prover_kernel.so (hand-written HIP for gfx950).
This package binds them with ctypes and mirrors the reference's host wrappers
(src/cuda/ntt.rs, src/cuda/msm.rs, src/cuda/bn254.rs) with the same argument
meaning.  torch is used only to allocate device memory and to name streams.
"""
from ._lib import HipError, lib, exported_symbols, LIB_PATH  # noqa: F401
