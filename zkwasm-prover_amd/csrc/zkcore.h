// Internal header shared by the HIP translation units of the kernel library.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "ec.h"
#include "zkwasm_prover_kernel.h"
#include "ff.h"

#define ZK_CHECK(x)                         \
    do {                                    \
        hipError_t zk_err_ = (x);           \
        if (zk_err_ != hipSuccess) return zk_err_; \
    } while (0)

namespace zk {

__host__ __device__ inline uint32_t log2_exact(uint64_t n) {
    uint32_t k = 0;
    while ((1ull << k) < n) k++;
    return k;
}
__host__ __device__ inline bool is_pow2(uint64_t n) { return n && !(n & (n - 1)); }

__device__ __forceinline__ uint32_t bitrev32(uint32_t x, uint32_t bits) {
    return bits ? (__builtin_bitreverse32(x) >> (32 - bits)) : 0u;
}

// cross-unit helpers (defined in ntt.hip)
hipError_t fill_zero(void *p, size_t bytes, hipStream_t stream);
hipError_t coset_scale(Fr *s, const Fr *cp, uint32_t cn, uint32_t n, hipStream_t stream);
hipError_t ntt_run(Fr *a, Fr *b, const Fr *omegas, const Fr *n_inv, uint32_t log_n, hipStream_t stream,
                   int *passes_out);
int ntt_plan(uint32_t log_n, uint32_t *degs);  // degs: room for 32 pass degrees

}  // namespace zk
