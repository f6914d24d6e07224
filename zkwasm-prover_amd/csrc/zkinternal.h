// Launchers shared between translation units of the library (plain C++
// declarations: host code can include this without the device headers).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zkwasm_prover_kernel.h"

namespace zk {

// zero-fill by kernel (no memset node in a captured graph) (ntt.hip)
hipError_t fill_zero(void *p, size_t bytes, hipStream_t stream);
// the NTT's radix plan for 2^log_n: pass count (degs: room for 32) (ntt.hip)
int ntt_plan(uint32_t log_n, uint32_t *degs);

// witness synthesis and helpers for the prover core (witness.hip)
hipError_t wit_u32_rows_to_fr(void *col, const uint32_t *v, uint32_t rows, hipStream_t s);
hipError_t wit_lookup_values(uint32_t *vals, uint32_t u, uint32_t tsize, uint64_t seed, hipStream_t s);
hipError_t wit_table_rows(void *col, uint32_t u, uint32_t tsize, hipStream_t s);
hipError_t wit_gate_rows(void *d, const void *a, const void *b, const void *c, uint32_t u, hipStream_t s);
hipError_t wit_mul_rows(void *d, const void *x, const void *y, uint32_t u, hipStream_t s);
hipError_t wit_indicator(void *out, uint32_t n, uint32_t lo, uint32_t hi, hipStream_t s);
hipError_t wit_zero_rows(void *col, uint32_t lo, uint32_t hi, hipStream_t s);
hipError_t wit_gather(void *dst, const void *src, const uint32_t *idx, uint32_t rows, hipStream_t s);
hipError_t wit_cycle_values(void *const *cols_dev, const uint32_t *head, uint32_t ncols, uint32_t u, hipStream_t s);
hipError_t wit_sigma_rows(void *const *sig_dev, const uint32_t *target, const void *deltas, const void *omegas,
                          uint32_t ncols, uint32_t n, uint32_t u, hipStream_t s);
hipError_t wit_fill_value(void *col, const void *v, uint32_t n, hipStream_t s);
hipError_t wit_any_nonzero(const void *col, uint64_t lo, uint64_t hi, uint32_t *flag, hipStream_t s);

// h = q (y^G (a b + c - d) + sum_g y^(G-1-g) (f_g a - e_g)) on the extended
// coset: the benchmark circuit's gates in one pass, G <= 32 (evalh.hip)
hipError_t gate_eval_h(void *h, const void *q, const void *a, const void *b, const void *c, const void *d,
                       const void *const *f, const void *const *e, int ng, const void *y, int n, hipStream_t stream);

// out = (x + beta)^-1 element-wise, temp: n Fr; and eval_logup_z_compose's
// steps after its in-place table inversion, given that inverse (inv_logup.hip)
hipError_t batch_invert_shifted(void *out, const void *x, const void *beta, void *temp, int n, hipStream_t stream);
hipError_t logup_z_inverted(void *z, void *input, const void *table_inv, const void *multiplicity, void *last_z,
                            int last_z_index, int n, hipStream_t stream);

// NTTs whose first pass reads src[m] (2^log_n elements, left unchanged)
// instead of buf[m]: the values of copying src into buf and running ntt with
// tmp.  swapped[m] set: the result is in tmp[m], else in buf[m]. (ntt.hip)
hipError_t ntt_batch_from(const void *const *src, void *const *buf, void *const *tmp, const void *omegas,
                          const void *n_inv, int log_n, int count, unsigned char *swapped, hipStream_t s);

// Commitment hook of a prover-core job: when set, the job's batched
// commitments go through fn instead of its own zk_msm_batch call.  fn gets
// the group's MSM items (inputs ready once `stream` reaches its current
// position) and returns their affine results in host memory (64 B each,
// zeros for the identity) after the MSMs have read their inputs.  The proof
// server sets it to merge the MSMs of concurrent proofs. (prover_core.cpp)
// One MSM through the pipeline with the device tail, enqueued on stream
// without host waits or host allocation (msm.hip)
hipError_t msm_one_async(void *out, const void *points, const void *scalars, uint32_t n, uint32_t c, void *scratch,
                         size_t scratch_bytes, hipStream_t stream);

// zk_msm_batch at the default window whose count results (Jacobian, 96 B
// each) are also returned in host memory, without a copy per item (msm.hip)
hipError_t msm_batch_host(void *out, const zk_msm_item *items, int count, void *scratch, size_t scratch_bytes,
                          hipStream_t stream, uint8_t *jac_host);

// out = sum_i v^(count-1-i) polys[i] (Horner per element; polys: host array
// of device pointers, none equal to out) (multiopen.hip)
hipError_t poly_horner_combine(void *out, const void *const *polys, int count, const void *v, int n,
                               hipStream_t stream);

typedef hipError_t (*CommitFn)(void *ctx, const void *items, int count, uint8_t *out_affine, hipStream_t stream);
hipError_t job_set_commit_fn(void *job, CommitFn fn, void *ctx);

}  // namespace zk
