/*
 * zkwasm_prover_kernel.h -- C ABI of the MI355X (gfx950) arithmetic core.
 *
 * Drop-in for the `extern "C"` block the reference prover links as the static
 * library `zkwasm_prover_kernel` (src/cuda/bn254_c.rs:4-309).  Every symbol of
 * that block that is on the KZG hot path is exported here under the same name
 * with the same argument order and meaning; `CUstream_st *` becomes
 * hipStream_t and `cudaError` becomes hipError_t (0 = success).  Each
 * declaration names the reference declaration it replaces.
 *
 * Conventions (as in the reference):
 *   - every buffer is a device pointer; Fr/Fp elements are 32 B, 4 x u64
 *     little-endian in Montgomery form (R = 2^256); G1 affine 64 B {x, y}
 *     (identity (0, 0)), G1 Jacobian 96 B {x, y, z} (identity z = 0);
 *   - scalars such as beta, gamma, y, n^-1 are passed by device pointer;
 *   - "arrays of buffers" are device arrays of device pointers;
 *   - sizes are element counts; n is a power of two unless stated otherwise;
 *   - calls are asynchronous on `stream` (0 = the null stream) and never
 *     allocate device memory; entries the reference runs on the default
 *     stream (no stream argument) run on the null stream here too.
 * Invalid arguments return hipErrorInvalidValue instead of asserting.
 */
#ifndef ZKWASM_PROVER_KERNEL_H
#define ZKWASM_PROVER_KERNEL_H

#include <hip/hip_runtime.h>
#include <stdbool.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- NTT (src/cuda/bn254_c.rs:6-16, :196, :31-39) --------------------- */

/* bn254_c.rs:6 `ntt`.  Natural-order DFT of buf (2^log_n Fr) with the
 * w^i table from expand_omega_buffer; n_inv != NULL scales by *n_inv (iNTT).
 * The result is in tmp when *swap is set on return, else in buf.  pq and
 * max_deg are accepted and unused (the in-pass twiddles come from omegas). */
hipError_t ntt(void *buf, void *tmp, const void *pq, const void *omegas, const void *n_inv, int log_n,
               int max_deg, bool *swap, hipStream_t stream);

/* bn254_c.rs:196 `expand_omega_buffer(buf, n, may_bit_reversed)`; the
 * second argument is log2 of the length.  buf[0] = 1, buf[1] = w on entry;
 * fills buf[i] = w^i for i < 2^log_n (natural order for every size). */
hipError_t expand_omega_buffer(void *buf, int log_n, int may_bit_reversed);

/* bn254_c.rs:31 `extended_prepare`.  to_coset == 0: zero [size,
 * extended_size) then s[i] *= coset_powers[(i % n) - 1] for i < size with
 * i % n != 0; to_coset != 0: the same scaling over [0, extended_size). */
hipError_t extended_prepare(void *s, const void *coset_powers, unsigned coset_powers_n, int size,
                            int extended_size, int to_coset, hipStream_t stream);

/* ---- MSM (src/cuda/bn254_c.rs:248-280) ---------------------------------- */

/* bn254_c.rs:261 `msm`.  Buffers per src/cuda/msm.rs:297-311: res = this
 * MSM's curve slice ((windows << window_bits) + 2 * workers Jacobian points),
 * four index buffers of array_len * windows u32, sort_buf of 3 * array_len *
 * windows u32.  They are used as scratch; the window is chosen internally.
 * Call batch_msm_collect afterwards for the result. */
hipError_t msm(void *res, const void *points, const void *scalars, void *indices_buf0, void *indices_buf1,
               void *indices_buf2, void *indices_buf3, void *sort_buf, void *acc_indices_buf, int array_len,
               int windows, int window_bits, int threads, int workers,
               int prepared_sort_indices_temp_storage_bytes, int skip_zero, hipStream_t stream);

/* bn254_c.rs:248 `batch_msm_collect`.  buckets: device array of batch_size
 * pointers to the curve slices given to msm(); afterwards each slice starts
 * with the MSM result, a Jacobian point with z = 1 (or the identity
 * (0, 1, 0)).  The remain arrays are unused.  The final window combination
 * runs on host cores and the call returns after the results are written
 * (the reference's caller synchronises right after it, src/cuda/msm.rs:436-439);
 * with ZK_MSM_TAIL=device in the environment it stays on the GPU and is
 * asynchronous.  A slice msm() did not finish (its header cleared or never
 * written) returns hipErrorInvalidValue; on the asynchronous device path that
 * slice's 96 result bytes are all 0xff instead (no point, not the identity). */
hipError_t batch_msm_collect(void *remain_indices, void *remain_acc, void *next_remain_indices,
                             void *next_remain_acc, void *buckets, unsigned workers, unsigned windows,
                             unsigned window_bits, unsigned batch_size, hipStream_t stream);

/* ---- batch inversion / logup grand sum (bn254_c.rs:149, :226, :238) ---- */

/* bn254_c.rs:149 `logup_sum_input_inv`: sum[i] = (init == 0) ? 1/(input[i]+beta)
 * : sum[i] + 1/(input[i]+beta).  temp: n Fr of scratch. */
hipError_t logup_sum_input_inv(void *sum, const void *input, void *temp, const void *beta, int init, int n,
                               hipStream_t stream);

/* bn254_c.rs:226 `eval_logup_z_compose`: z[i] = input[i] - m[i]/(table[i]+beta)
 * then the exclusive scan z[i] = last_z[0] + sum_{j<i} z[j]; finally
 * last_z[0] = z[last_z_index].  input and table are scratch afterwards. */
hipError_t eval_logup_z_compose(void *z, void *input, void *table, const void *multiplicity, const void *beta,
                                void *last_z, int last_z_index, int n, hipStream_t stream);

/* bn254_c.rs:238 `eval_logup_z_pure`: in-place exclusive scan of z with carry
 * last_z[0]; last_z[0] = z[last_z_index] afterwards.  input and table are
 * scratch (n Fr each). */
hipError_t eval_logup_z_pure(void *z, void *input, void *table, void *last_z, int last_z_index, int n,
                             hipStream_t stream);

/* ---- lookup multiplicities m(X) (bn254_c.rs:281, :295) ----------------- */

/* bn254_c.rs:281 `prepare_table_lookup`: indexes the usable rows
 * [0, unusable_rows_start) of table (n = 2^k Fr) for calc_m.  Every buffer
 * argument is n Fr as the reference's caller passes (src/lib.rs:407-412);
 * start_offset receives the index (a hash table of 8n u32 slots) and must be
 * kept for the calc_m calls of this table; the other buffers are unused.
 * Returns hipErrorInvalidValue for k outside [1, 27] or
 * unusable_rows_start > n. */
hipError_t prepare_table_lookup(void *table, void *table_lowest_u32, void *sorted_table_lowest_u32, void *index,
                                void *sorted_index, void *start_offset, void *candidate_sort_temp_storage,
                                unsigned candidate_sort_temp_storage_bytes, unsigned k, unsigned unusable_rows_start,
                                hipStream_t stream);

/* bn254_c.rs:295 `calc_m`: for each usable input row i, the matched table row
 * j is the largest usable row with table[j] == input[i] (Montgomery words
 * compared); m[j] += (number of inputs matched to j).  m accumulates across
 * calls (the caller zeroes it once per lookup).  matched_index (n Fr) is
 * scratch; the first u32 of sorted_matched_index is set to 1 when some usable
 * input has no usable table row equal to it (the reference asserts), else 0. */
hipError_t calc_m(void *m, void *table, void *input, const void *sorted_index, const void *start_offset,
                  void *matched_index, void *sorted_matched_index, void *candidate_sort_temp_storage,
                  unsigned candidate_sort_temp_storage_bytes, unsigned k, unsigned unusable_rows_start,
                  hipStream_t stream);

/* ---- eval_h evaluator and argument kernels ---------------------------- */

/* bn254_c.rs:18 `field_op`: res[i] = L op R with L = l[(i+l_rot) mod n]
 * (* l_c) or l_c when l is NULL, R likewise; op 0 add, 1 mul, 2 copy L,
 * 3 sub.  Negative rotations wrap (mod n). */
hipError_t field_op(void *res, const void *l, int l_rot, const void *l_c, const void *r, int r_rot,
                    const void *r_c, int size, int op, hipStream_t stream);

/* bn254_c.rs:81 `field_op_batch_mul_sum`: v = [c0, a00, a01, NULL, c1, a10, NULL,
 * ...] (v_n entries), rot = one i32 per column pointer;
 * res[i] += sum_t c_t * prod_j v_tj[(i + rot_tj) & (n-1)]. */
hipError_t field_op_batch_mul_sum(void *res, const void *v, const void *rot, int v_n, int n,
                                  hipStream_t stream);

/* bn254_c.rs:198 `field_mul_zip`: buf[i] *= coeff[i % coeff_n]. */
hipError_t field_mul_zip(void *buf, const void *coeff, int coeff_n, int n);

/* bn254_c.rs:71 `field_sum` (declared, never called by the reference):
 * res[i] = sum_j v[j][i] * omegas[(rot_j * i) mod n] (* *v_c[j] when v_c[j]
 * is non-null), i.e. coefficient vectors rotated by w^rot_j and summed.  v,
 * v_c: device arrays of v_n device pointers; v_rot: device array of v_n i32;
 * omegas: w^i for i < n; v_n <= 256.  Null stream. */
hipError_t field_sum(void *res, const void *v, const void *v_c, const void *v_rot, const void *omegas, int v_n,
                     int n);

/* bn254_c.rs:90 `lookup_eval_h` (declared, never called by the reference):
 * the permuted-column lookup's folds res = ((((res*y + l0 (1 - z))*y
 * + l_last (z^2 - z))*y + l_active (z(wX)(a' + beta)(s' + gamma)
 * - z input table))*y + l0 (a' - s'))*y + l_active (a' - s')(a' - a'(w^-1 X)),
 * with input / table the compressed expressions including beta / gamma and
 * w X at index i + rot. */
hipError_t lookup_eval_h(void *res, const void *input, const void *table, const void *permuted_input,
                         const void *permuted_table, const void *z, const void *l0, const void *l_last,
                         const void *l_active_row, const void *y, const void *beta, const void *gamma, int rot,
                         int n, hipStream_t stream);

/* bn254_c.rs:41 `permutation_eval_h_p1`: res = (res*y + l0*(1 - first))*y
 * + l_last*(last^2 - last). */
hipError_t permutation_eval_h_p1(void *res, const void *first_set, const void *last_set, const void *l0,
                                 const void *l_last, const void *y, int n);

/* bn254_c.rs:51 `permutation_eval_h_p2`: for j in 1..n_set:
 * res = res*y + l0*(set[j][i] - set[j-1][(i+rot) & (n-1)]). */
hipError_t permutation_eval_h_p2(void *res, const void *set, const void *l0, const void *l_last, const void *y,
                                 int n_set, int rot, int n);

/* bn254_c.rs:62 `permutation_eval_h_l`: res[i] += beta*p[i] (+ gamma at i=0). */
hipError_t permutation_eval_h_l(void *res, const void *beta, const void *gamma, const void *p, int n,
                                hipStream_t stream);

/* bn254_c.rs:108 `logup_eval_h` (cuda/bn254.cu:867-918 formula). */
hipError_t logup_eval_h(void *res, const void *input_product, const void *input_product_sum, const void *table,
                        const void *multiplicity, const void *z_first, const void *z_last, const void *l0,
                        const void *l_last, const void *l_active_row, const void *y, int rot, int n,
                        hipStream_t stream);

/* bn254_c.rs:125 `logup_eval_h_extra_inputs` (cuda/bn254.cu:920-949). */
hipError_t logup_eval_h_extra_inputs(void *res, const void *input_product, const void *input_product_sum,
                                     const void *z, const void *l_active_row, const void *y, int rot, int n,
                                     hipStream_t stream);

/* bn254_c.rs:137 `logup_eval_h_z_set`: permutation_eval_h_p2 on `stream`. */
hipError_t logup_eval_h_z_set(void *res, const void *set, const void *l0, const void *l_last, const void *y,
                              int n_set, int rot, int n, hipStream_t stream);

/* bn254_c.rs:159 `logup_eval_h_inputs_product_sum`: product = prod_j set[j],
 * product_sum = sum_j prod_{k != j} set[k]. */
hipError_t logup_eval_h_inputs_product_sum(void *product, void *product_sum, const void *sets, int n_set, int n,
                                           hipStream_t stream);

/* bn254_c.rs:168 `shuffle_eval_h` (cuda/bn254.cu:975-1016). */
hipError_t shuffle_eval_h(void *res, const void *input, const void *table, const void *z, const void *l0,
                          const void *l_last, const void *l_active_row, const void *y, int rot, int n);

/* bn254_c.rs:181 `shuffle_eval_h_v2` (cuda/bn254.cu:1018-1071). */
hipError_t shuffle_eval_h_v2(void *res, const void *inputs, const void *tables, const void *betas, int group_len,
                             const void *z, const void *l0, const void *l_last, const void *l_active_row,
                             const void *y, int rot, int n);

/* ---- multi-open (bn254_c.rs:200, :208, :218) ---------------------------- */

/* bn254_c.rs:200 `poly_eval`: res[0] = p(x) = sum_i p[i] x^i for the n
 * coefficients p, with x = x_table[0] (the caller's table holds x^(2^j),
 * src/eval_poly.rs:44-53; only its first entry is read).  res (n Fr) is
 * scratch; any n >= 1 (the reference needs a power of two and leaves res
 * unwritten for n = 1). */
hipError_t poly_eval(const void *p, void *res, const void *x, int n, hipStream_t stream);
/* Many polynomials at one point (replaces the per-polynomial poly_eval loop of
 * batch_poly_eval, src/eval_poly.rs:18-140): polys is a HOST array of count
 * device pointers (n coefficients each), out a HOST array of count device
 * pointers, each receiving one Fr, p_m(x); x is a device Fr.  scratch holds
 * zk_poly_eval_batch_scratch_bytes(n, count) bytes.  Asynchronous. */
size_t zk_poly_eval_batch_scratch_bytes(int n, int count);
hipError_t zk_poly_eval_batch(const void *const *polys, void *const *out, int count, const void *x, int n,
                              void *scratch, hipStream_t stream);

/* bn254_c.rs:208 `shplonk_h_x_merge`: res[i] = res[i]*v
 * + values[i] * prod_{j < diff_points_n} (omegas[i] - diff_points[j]).
 * Legacy default stream, as the reference. */
hipError_t shplonk_h_x_merge(void *res, const void *v, const void *values, const void *omegas,
                             const void *diff_points, int diff_points_n, int n);

/* bn254_c.rs:218 `shplonk_h_x_div_points`: values[i] /= prod_{j < points_n}
 * (omegas[i] - points[j]), points_n >= 1.  A zero product (omegas[i] equal to
 * a point; the reference asserts) leaves values[i] = 0.  Legacy default
 * stream. */
hipError_t shplonk_h_x_div_points(void *values, const void *omegas, const void *points, int points_n, int n);

/* ---- extensions: no counterpart in the reference's C ABI --------------- */

/* Division by (X - z) of the n coefficients in `in` (src/multiopen.rs:145-159
 * for GWC, :543-553 for Shplonk, done there on the host): out[i] = q_i for
 * i <= n-2 with in(X) = (X - z) q(X) + in(z), and out[n-1] = 0 (n = 1 copies
 * the single entry, as the reference's loop does).  out may equal in; z is a
 * device Fr; eval (device Fr, may be NULL) receives in(z).  temp:
 * zk_kate_divide_scratch_bytes(n) bytes of device memory. */
size_t zk_kate_divide_scratch_bytes(int n);
hipError_t zk_kate_divide(void *out, const void *in, const void *z, void *eval, int n, void *temp,
                          hipStream_t stream);
hipError_t zk_shplonk_h_x_merge_on(void *res, const void *v, const void *values, const void *omegas,
                                   const void *diff_points, int diff_points_n, int n, hipStream_t stream);
hipError_t zk_shplonk_h_x_div_points_on(void *values, const void *omegas, const void *points, int points_n, int n,
                                        hipStream_t stream);

/* Stream-taking forms of the reference entries that run on the legacy
 * default stream (expand_omega_buffer, field_mul_zip, permutation_eval_h_p1,
 * permutation_eval_h_p2, shuffle_eval_h, shuffle_eval_h_v2): same arguments
 * plus the stream, so concurrent provers on their own streams overlap. */
hipError_t zk_expand_omega_buffer_on(void *buf, int log_n, hipStream_t stream);
hipError_t zk_field_mul_zip_on(void *buf, const void *coeff, int coeff_n, int n, hipStream_t stream);
hipError_t zk_permutation_eval_h_p1_on(void *res, const void *first_set, const void *last_set, const void *l0,
                                       const void *l_last, const void *y, int n, hipStream_t stream);
hipError_t zk_permutation_eval_h_p2_on(void *res, const void *set, const void *l0, const void *l_last,
                                       const void *y, int n_set, int rot, int n, hipStream_t stream);
hipError_t zk_shuffle_eval_h_on(void *res, const void *input, const void *table, const void *z, const void *l0,
                                const void *l_last, const void *l_active_row, const void *y, int rot, int n,
                                hipStream_t stream);
hipError_t zk_shuffle_eval_h_v2_on(void *res, const void *inputs, const void *tables, const void *betas,
                                   int group_len, const void *z, const void *l0, const void *l_last,
                                   const void *l_active_row, const void *y, int rot, int n, hipStream_t stream);

/* One permutation set's term of h on the extended coset in one pass
 * (replaces the field_op / permutation_eval_h_l / extended-NTT sequence of
 * src/eval_h.rs:700-840): h = h*y + l_active * (z(wX) prod_c (v_c + beta
 * sigma_c + gamma) - z(X) prod_c (v_c + gamma + bd_c X)); v, sigma: device
 * arrays of nc <= 16 coset columns, bd: nc Fr (beta delta^c), X: coset
 * points, rot: one row's shift on the extended domain. */
hipError_t zk_permutation_eval_h_set(void *h, const void *z, const void *v, const void *sigma, int nc,
                                     const void *beta, const void *gamma, const void *bd, const void *X,
                                     const void *l_active, const void *y, int rot, int n, hipStream_t stream);

/* Batch inversion in place, zeros stay zero (halo2 BatchInvert, used at
 * src/permutation.rs:73-75 and src/shuffle.rs:163-165).  temp: n Fr. */
hipError_t zk_batch_invert(void *a, void *temp, int n, hipStream_t stream);

/* One permutation set's grand product on the device, replacing the CPU
 * src/permutation.rs:44-128 for one chunk of columns: values/sigmas are
 * device arrays of ncols column pointers; delta_start = DELTA^(set*chunk_len).
 * z[i] = exclusive running product of
 *   prod_j (v_j + beta*delta_start*DELTA^j*omega^i + gamma) / prod_j (v_j + beta*sigma_j + gamma).
 * temp: n Fr of scratch. */
hipError_t zk_permutation_z_set(void *z, const void *values, const void *sigmas, int ncols, const void *beta,
                                const void *gamma, const void *delta_start, const void *delta,
                                const void *omega, void *temp, int n, hipStream_t stream);

/* Chains the sets (src/permutation.rs:96-147): zs is a device array of nsets
 * pointers; z_s[0..=u] *= prod_{s'<s} z_{s'}[u] (tails taken before
 * chaining).  tails: nsets Fr of scratch. */
hipError_t zk_permutation_chain(void *zs, int nsets, int u, int n, void *tails, hipStream_t stream);

/* Shuffle grand product (src/shuffle.rs:142-209): inputs/shuffles are device
 * arrays of group_len pointers, betas[j] = beta^(j+1); z = exclusive running
 * product of prod_j (input_j + betas[j]) / prod_j (shuffle_j + betas[j]).
 * temp: n Fr. */
hipError_t zk_shuffle_z(void *z, const void *inputs, const void *shuffles, int group_len, const void *betas,
                        void *temp, int n, hipStream_t stream);

/* Batched NTT: count independent transforms (sizes may differ) with one
 * launch per pass for the whole batch instead of one `ntt` call per
 * polynomial (src/cuda/ntt.rs:61-93).  Each item is an `ntt` call's
 * arguments; items is a HOST array; swapped[m] (host) is set when result m
 * landed in tmp.  Asynchronous on stream. */
typedef struct {
    void *buf;
    void *tmp;
    const void *omegas;  /* expand_omega_buffer table of 2^log_n entries */
    const void *n_inv;   /* null, or n^-1 (iNTT) */
    int log_n;
    int reserved;
} zk_ntt_item;
hipError_t zk_ntt_batch(const zk_ntt_item *items, int count, unsigned char *swapped, hipStream_t stream);

/* Coset extension in one pipeline (the reference's extended_prepare followed
 * by ntt, bn254_c.rs:31 and :6, as called from src/eval_h.rs:250-337): for
 * each item, src's 2^log_n coefficients zero-padded to 2^(log_n + log_ext),
 * scaled by coset_powers[(i % coset_powers_n) - 1] where i % coset_powers_n
 * != 0, and transformed with omegas (the 2^(log_n + log_ext) table).  The
 * first pass reads only the coefficients and scales them on load; no zero
 * padding is written.  Results equal the three-call sequence; each lands in
 * buf, or in tmp when swapped[m] is set.  src must not alias buf or tmp. */
typedef struct {
    const void *src;
    void *buf;
    void *tmp;
    const void *omegas;
    int log_n;
    int log_ext;
} zk_coset_item;
hipError_t zk_coset_extend_batch(const zk_coset_item *items, int count, const void *coset_powers,
                                 int coset_powers_n, unsigned char *swapped, hipStream_t stream);

/* Radix choice per size (autotuner hook): largest pass degree for 2^log_n
 * transforms, 1..8 (0 = default 8). */
hipError_t zk_ntt_set_max_deg(int log_n, int max_deg);
int zk_ntt_get_max_deg(int log_n);
/* Autotuner: times each distinct radix plan (largest pass degree 4..8) for
 * 2^log_n on the given buffers (overwritten) and keeps the fastest; ms
 * (optional, 9 floats) gets the time per max degree.  The table is safe to
 * read and set from any thread; an autotune is a set-up call: NTTs of its
 * size run meanwhile on other threads stay correct but use its trial plans. */
hipError_t zk_ntt_autotune(void *buf, void *tmp, const void *omegas, int log_n, int reps, hipStream_t stream,
                           int *best, float *ms);


/* MSM with caller-owned scratch of zk_msm_scratch_bytes(n, c) bytes; the
 * result (Jacobian, z = 1) lands in out[0..96); out must hold
 * ZK_MSM_OUT_BYTES bytes.  c = 0 picks the default window.  Synchronous. */
#define ZK_MSM_OUT_BYTES (64 * 1024)
size_t zk_msm_scratch_bytes(int n, int c);
hipError_t zk_msm(void *out, const void *points, const void *scalars, int n, int c, void *scratch,
                  size_t scratch_bytes, hipStream_t stream);

/* zk_msm without any allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The pipeline's scratch is a
 * device array of ZK_MSM_NOALLOC_SCRATCH_BYTES inside the library (one per
 * GPU, held until the process ends); n up to ZK_MSM_NOALLOC_MAX_N fits at the
 * default window (c = 0), and sizes or windows whose scratch exceeds the
 * array return hipErrorInvalidValue.  Asynchronous: the call only enqueues
 * kernels on stream (window combination on the device); the result
 * (Jacobian, z = 1) lands in out[0..96) when the stream reaches it, 96 bytes
 * of 0xff if the MSM could not finish; out holds ZK_MSM_OUT_BYTES.  Every call
 * on one GPU shares its array: calls that overlap on one GPU (other streams
 * or threads) race, so serialise them, e.g. on one stream. */
#define ZK_MSM_NOALLOC_SCRATCH_BYTES (40u << 20)
#define ZK_MSM_NOALLOC_MAX_N (1 << 17)
hipError_t zk_msm_noalloc(void *out, const void *points, const void *scalars, int n, int c, hipStream_t stream);

/* Batched MSM: count independent MSMs (device point / scalar arrays, any
 * sizes; items may share points) through one pass of the pipeline, one
 * launch per phase for the whole batch.  Replaces the per-polynomial loop of
 * the reference's batch_msm driver (src/cuda/msm.rs:369-455), whose MSMs go
 * one by one over two streams.  items is a HOST array; result m (Jacobian,
 * z = 1) lands in out + m * ZK_MSM_OUT_BYTES.  Synchronous.
 * table_c = 0: points holds n affine points.  table_c = c in [8, 20]: points
 * is a fixed-base table that zk_msm_table_build made from n points for window
 * c (the argument c of the batch call then does not apply to the item).  The
 * plain items, and the table items of each window, go through one pass each. */
typedef struct {
    const void *points;
    const void *scalars;
    int n;
    int table_c;
} zk_msm_item;
size_t zk_msm_batch_scratch_bytes(const zk_msm_item *items, int count, int c);
hipError_t zk_msm_batch(void *out, const zk_msm_item *items, int count, int c, void *scratch,
                        size_t scratch_bytes, hipStream_t stream);
/* The same without host waits: the window combination runs on the device,
 * so the call only enqueues kernels (capturable into a HIP graph).  An item
 * whose MSM did not finish reads as 96 bytes of 0xff (no point). */
hipError_t zk_msm_batch_async(void *out, const zk_msm_item *items, int count, int c, void *scratch,
                              size_t scratch_bytes, hipStream_t stream);

/* Fixed-base MSM tables.  No counterpart in the reference, whose msm()
 * takes the bases afresh on every call: KZG commitments always use the same
 * SRS points, so a table of T[w * n + i] = 2^(c w) P_i for the ceil(255 / c)
 * windows w (affine, 64 B each, the identity as (0, 0)) lets every window's
 * digits share ONE set of 2^(c-1) buckets: the buckets are summed once
 * instead of once per window, so a wider window pays (c = 20 at 2^22 points:
 * 13 bucket additions per scalar instead of 15).  Results are the same points
 * as plain MSMs over P.  zk_msm_table_bytes: the table's size for n points at
 * window c (8 <= c <= 20; 0 if invalid); zk_msm_table_window: the window with
 * the fewest additions for n points; zk_msm_table_build: fills table from the
 * n device points (asynchronous; set-up work, about 0.1 s at 2^22 points).
 * Use a table as a zk_msm_item with table_c = c. */
size_t zk_msm_table_bytes(int n, int c);
int zk_msm_table_window(int n);
hipError_t zk_msm_table_build(void *table, const void *points, int n, int c, hipStream_t stream);

/* HIP graph capture / replay of a launch sequence (no reference counterpart:
 * the reference re-issues every launch per proof, src/eval_h.rs, src/lib.rs).
 * Between begin and end only asynchronous entries may be called on stream
 * (not the legacy default stream); exec is an opaque executable graph. */
hipError_t zk_graph_begin(hipStream_t stream);
hipError_t zk_graph_end(hipStream_t stream, void **exec);
hipError_t zk_graph_launch(void *exec, hipStream_t stream);
hipError_t zk_graph_destroy(void *exec);

/* MSM phase timing for the bench: HIP events on the launch stream around each
 * phase (count, scan, scatter, accumulate, fixup, reduce, finalize); reading
 * waits for the profiled calls' events, sums them and returns the number of
 * profiled calls (-1 on a HIP error).  A profiled call does not wait for its
 * events (they go to a ring of 64 sets, drained when read or when full); the
 * state is global: enable it only while one thread runs MSMs. */
void zk_msm_profile_enable(int on);
int zk_msm_profile_read(double *times, int ntimes, int reset);
int zk_msm_default_window(int n);
/* Window table per size class ceil(log2 n) (0 = model default), and the
 * autotuner that fills it: times zk_msm for windows around the model's choice
 * on the given device points / scalars and keeps the fastest for every later
 * MSM of that class (msm, zk_msm, zk_msm_batch).  ms: optional 18 floats.
 * Thread-safety as for the NTT table above. */
/* The window msm() picks with the reference driver's buffers (sizes of
 * src/cuda/msm.rs:284-311) for array_len points, 0 if they cannot hold its
 * scratch.  Host-only. */
int zk_msm_plan_window(int array_len, int windows, int window_bits, int workers);
hipError_t zk_msm_set_window(int log_n, int c);
int zk_msm_get_window(int log_n);
hipError_t zk_msm_autotune(const void *points, const void *scalars, int n, int reps, hipStream_t stream, int *best,
                           float *ms);

/* Element-wise field / curve operations and calibration kernels (tests and
 * bench only).  zk_fr_op / zk_fp_op: op 0 a+b, 1 a*b, 2 a-b, 3 a^-1 (0 -> 0),
 * 4 a^2, 5 -a, 6 to Montgomery, 7 from Montgomery; b may be null (read as 0).
 * zk_g1_op (a, b affine, out Jacobian): op 0 mixed XYZZ add, 1 Jacobian add,
 * 2 XYZZ double of a, 3 Jacobian double of a, 4 XYZZ add, normalised; b is
 * read by every op.  Out-of-range ops, null buffers and empty or malformed
 * calibration launches (block outside 1..1024, threads not a multiple of it,
 * iters < 0) return hipErrorInvalidValue. */
hipError_t zk_fr_op(void *out, const void *a, const void *b, int op, int n, hipStream_t stream);
hipError_t zk_fp_op(void *out, const void *a, const void *b, int op, int n, hipStream_t stream);
hipError_t zk_g1_op(void *out, const void *a, const void *b, int op, int n, hipStream_t stream);
/* the 29-bit-limb Fp of the bucket accumulation through its conversions:
 * op 0 a*b, 1 a+b, 2 a-b, 3 a^2, 4 round trip (32-bit Montgomery in and out) */
hipError_t zk_fp29_op(void *out, const void *a, const void *b, int op, int n, hipStream_t stream);
/* 29-bit-limb multiply throughput (calibration): iters x 2 products per thread */
hipError_t zk_bench_mul29(void *buf, int threads, int block, int iters, int variant, hipStream_t stream);
hipError_t zk_bench_mul(void *buf, int threads, int block, int iters, hipStream_t stream);
hipError_t zk_bench_mul_cios(void *buf, int threads, int block, int iters, hipStream_t stream);
hipError_t zk_bench_dbl(void *buf, int threads, int block, int iters, hipStream_t stream);
hipError_t zk_bench_madd(void *acc, const void *pts, int npts, int threads, int block, int iters,
                         hipStream_t stream);
hipError_t zk_copy(void *dst, const void *src, size_t bytes, hipStream_t stream);
/* seeded synthetic inputs (bench): uniform Fr (Montgomery) and points k_i * G */
hipError_t zk_rand_fr(void *out, int n, unsigned long long seed, hipStream_t stream);
hipError_t zk_gen_points(void *out, int n, unsigned long long seed, hipStream_t stream);
/* out[i] = s[i] G (affine, G = (1, 2)) for n Montgomery scalars s */
hipError_t zk_g1_mul_scalars(void *out, const void *scalars, int n, hipStream_t stream);

/* ---- host-side mirror of the reference's MSM driver -------------------- */
/* src/cuda/msm.rs:274-283 window policy */
void zkp_msm_policy(int count, int len, int *window_bits, int *windows, int *workers);
/* src/cuda/msm.rs:239-455 batch_msm over device-resident scalar sets;
 * out_affine: count affine points in host memory. */
hipError_t zkp_batch_msm(void *out_affine, const void *points_dev, const void *const *scalars_dev, int count,
                         int len, int skip_zero);

/* ---- host-side mirror of the vanishing construction ------------------- */
/* src/eval_h.rs:464-540: h (2^log_ext Fr, device) *= t_evals[i mod t_len];
 * extended iNTT (ext_intt_omegas, ext_divisor); coset un-distribution with
 * coset_powers = [g_coset_inv, g_coset]; one commitment per quotient piece of
 * 2^log_n coefficients against g (device, 2^log_n affine).  commitments
 * (host): quotient_degree affine points.  *coeffs = h or tmp, whichever holds
 * the coefficients.  batched = 0: the reference driver's MSMs (zkp_batch_msm);
 * 1: one zk_msm_batch pass. */
hipError_t zkp_vanishing_commit(void *h, void *tmp, const void *t_evals, int t_len, const void *ext_intt_omegas,
                                const void *ext_divisor, const void *coset_powers, int log_n, int log_ext,
                                int quotient_degree, const void *g, void *commitments, void **coeffs, int batched);
/* src/eval_h.rs:546-583: h_pieces (host, 2^log_n Fr) = sum_i piece_i xn^i by
 * Horner, folded in place into the last piece of coeffs; xn: device x^n. */
hipError_t zkp_vanishing_pieces(void *coeffs, const void *xn, int log_n, int quotient_degree, void *h_pieces);

/* ---- Fiat-Shamir transcript (host) -------------------------------------- */
/* halo2_proofs' Blake2bWrite, which the reference's create_proof drives
 * (src/lib.rs:205-1194; not part of its CUDA ABI): Blake2b-512 personalised
 * "Halo2-Transcript"; a point absorbs 0x01 || x || y and a scalar 0x02 || s
 * (canonical little-endian 32-byte values); a challenge absorbs 0x00, and is
 * the 64-byte digest of the state so far read little-endian mod r
 * (from_bytes_wide), returned in Montgomery form.  Host functions: no device
 * is touched. */
int zk_blake2b(void *out, int outlen, const void *in, size_t inlen, const void *personal16);
void *zk_transcript_new(void);
void zk_transcript_free(void *t);
int zk_transcript_common_point(void *t, const void *x, const void *y);
int zk_transcript_common_scalar(void *t, const void *s);
int zk_transcript_squeeze(void *t, void *mont_out, void *digest64);
/* halo2's Fr::from_bytes_wide: 64 little-endian bytes mod r, Montgomery out */
int zk_fr_from_bytes_wide(const void *in64, void *mont_out);

/* ---- prover core: one proof's device arithmetic ------------------------ */
/* The PLONKish benchmark circuit: a*b + c - d = 0 under a selector q (fixed
 * column 0) on the usable rows; `fixed` - 2 coefficient gates
 * q (f_j a - e_j) = 0 over the fixed columns 2.. (f_j) and advice columns e_j;
 * `lookups` advice columns range-checked against a fixed table (fixed column
 * 1) of 2^table_log values; a permutation over `perm_cols` columns -- the
 * `instance` instance columns first, then perm_cols - instance advice
 * columns -- joined by random copy cycles; `shuffles` column pairs, one a row
 * permutation of the other on 2^shuffle_log rows; `blinding` random rows at
 * the end.  Needs fixed >= 2, 0 <= instance <= perm_cols and advice >=
 * 4 + lookups + (perm_cols - instance) + 2 * shuffles + (fixed - 2); other
 * shapes are rejected with hipErrorInvalidValue. */
typedef struct {
    int k, advice, fixed, instance, lookups, table_log, perm_cols, shuffles, shuffle_log, blinding;
} zkp_circuit;

typedef struct {
    int ncommit;         /* commitments written */
    int z_ok;            /* every grand product closed at row u */
    int quotient_ok;     /* h(X) has degree < 3n: all constraints hold */
    int z_fail;          /* when !z_ok: bit 0 logup, bit 1 permutation, bit 2 shuffle */
    int open_ok;         /* GWC: each batch evaluates at its point to the
                            combination of the evaluations; Shplonk: L(u) = 0 */
    int nevals;          /* polynomial evaluations of the opening */
    double phase_ms[7];  /* witness, advice commit + iNTT, multiplicities,
                            grand products, h evaluation, vanishing, opening */
    double total_ms;     /* phases 1-6 */
    unsigned char transcript[32]; /* a last challenge's digest: fingerprint of
                                     every commitment and evaluation */
    unsigned long long challenges[4][4]; /* beta, gamma, y, x (Montgomery) */
} zkp_proof_info;

/* Device bytes of the resident key data and of one in-flight proof. */
size_t zkp_pk_bytes(const zkp_circuit *c);
size_t zkp_job_bytes(const zkp_circuit *c);
int zkp_commitment_count(const zkp_circuit *c);
/* Key data (SRS stand-ins, fixed columns, sigma, tables): built once per
 * circuit shape, shared by every job of it. */
hipError_t zkp_pk_create(const zkp_circuit *c, unsigned long long seed, hipStream_t stream, void **pk);
hipError_t zkp_pk_destroy(void *pk);
/* The key's SRS is a KZG setup for a secret tau drawn from the key seed:
 * g[i] = tau^i G (coefficient basis), g_lag[i] = L_i(tau) G (Lagrange basis).
 * zkp_pk_srs_secret copies tau (Montgomery Fr) -- a synthetic setup whose
 * trapdoor tests use to check openings without pairings.
 * zkp_pk_fixed_commitments writes the verifying key's commitments (affine,
 * 64 B each) of the fixed columns q, t, f_0.. and the perm_cols sigma
 * columns; returns their count (fixed + perm_cols) or -1 if out is too small.
 * zkp_pk_vk_repr writes the key's transcript representation (Montgomery Fr):
 * Blake2b-512 "Halo2-Verify-Key" over the key's canonical description (domain,
 * column and argument counts, every commitment) read by from_bytes_wide, the
 * first scalar each proof absorbs (vk.hash_into, src/lib.rs:242). */
hipError_t zkp_pk_srs_secret(void *pk, void *tau_out);
int zkp_pk_fixed_commitments(void *pk, void *out, int max);
hipError_t zkp_pk_vk_repr(void *pk, void *out32);
/* Per-proof device state (witness, grand products, extended polys, h). */
hipError_t zkp_job_create(void *pk, void **job);
hipError_t zkp_job_destroy(void *job);
/* One proof's arithmetic (src/lib.rs:205-1194 phase order, then the
 * multi-open of src/multiopen.rs) for seed, on stream; commitments (host,
 * room for zkp_commitment_count affine points) gets info.ncommit points in
 * transcript order: instance (absorbed, not written), advice, m, permutation
 * z, logup z, shuffle z, the vanishing argument's random polynomial, the
 * quotient pieces, then the opening's (GWC: one per rotation in the order
 * last, cur, next; Shplonk: h, L); h_pieces (host, 2^k Fr, may be null) the
 * folded quotient.  flags:
 * ZKP_BATCHED commits through zk_msm_batch (else one zk_msm per polynomial);
 * ZKP_TAMPER_GATE breaks the gate on one usable row (a self-test: the
 * quotient check must then fail); ZKP_SHPLONK opens with Shplonk
 * (create_proof_from_advices_with_shplonk), else GWC (_with_gwc);
 * ZKP_NO_OPEN stops after the quotient commitments; ZKP_TAMPER_EVAL moves
 * one opening evaluation off by one (a self-test: open_ok must then be 0). */
#define ZKP_BATCHED 1
#define ZKP_TAMPER_GATE 2
#define ZKP_SHPLONK 4
#define ZKP_NO_OPEN 8
#define ZKP_TAMPER_EVAL 16
/* ZKP_HOST_WITNESS: the advice and instance columns come from host memory
 * (zkp_set_host_witness) and are uploaded during the commitment phase,
 * overlapped with the MSMs, instead of being synthesised on the device. */
#define ZKP_HOST_WITNESS 32
/* ZKP_GRAPH: the h(X) round (coset extensions and constraint kernels) is
 * captured into a HIP graph on the job's first proof with this flag and
 * replayed on later ones (same results as the eager launches). */
#define ZKP_GRAPH 64
/* The last proof's opening: its evaluations (host, Montgomery Fr) in the
 * reference's order (src/lib.rs:1025-1078): h(x) -- computed, not written --
 * then instance, advice, fixed (q, t, f_j) and the random polynomial at x,
 * sigmas at x, each permutation z at x, xw and (all but the last) x w^-(b+1),
 * each lookup's m at x and z at x, xw, each shuffle z at x, xw; and the
 * opening challenges [v, y', u] (y', u zero for GWC).  Returns the number of
 * evaluations, or -1 if max is too small.
 * zkp_proof_bytes: the last proof's bytes (written points, compressed to 32
 * bytes -- canonical x little-endian, the parity of y in the top bit, the
 * identity all zeros -- and written scalars, canonical 32 bytes) in
 * transcript order; returns the length, -1 if max is too small; out = null
 * returns the length only.  zkp_proof_instance: the last proof's instance
 * columns (n Fr each, `rows` Fr apart in out). */
int zkp_proof_openings(void *job, void *evals, int max, void *challenges3);
/* Host-resident witness for the next ZKP_HOST_WITNESS proof, as
 * create_proof_from_advices receives it (src/lib.rs:205-257): advice_count =
 * advice and instance_count = instance columns of 2^k Fr (Lagrange values,
 * Montgomery) in host memory; the pointers are kept, not copied, until that
 * proof returns; pinned memory (zk_pinned_alloc) makes the uploads
 * asynchronous.  zkp_proof_witness copies the last proof's witness (advice,
 * instance) to host memory. */
hipError_t zkp_set_host_witness(void *job, const void *const *advice, int advice_count, const void *const *instance,
                                int instance_count);
hipError_t zkp_proof_witness(void *job, void *const *advice, void *const *instance);
long long zkp_proof_bytes(void *job, void *out, long long max);
hipError_t zkp_proof_instance(void *job, void *out, int rows);
hipError_t zkp_prove(void *job, unsigned long long seed, int flags, void *commitments, void *h_pieces,
                     zkp_proof_info *info, hipStream_t stream);

/* ---- proof server: continuous admission by HBM footprint -------------- */
/* Keys of every shape stay resident; `slots` workers (one stream each) start
 * queued requests FIFO as slots free up.  The per-proof device states are
 * created with the server (one per shape, then up to `slots` per shape,
 * smallest shapes first, within hbm_budget; 0: free memory less 4 GiB) and
 * reused, so serving allocates nothing; a request whose shape has no idle
 * state lets others pass until it has been passed 8 times.  (If one state of
 * every shape does not fit the budget, states are created on demand and idle
 * ones of other shapes freed.)  A proof joins others in flight only when more
 * than two requests wait per proof running (the environment variable
 * ZKS_OVERLAP_RATIO, 1..64, sets that number; ZKS_EAGER: whenever a slot and
 * a state are free).  flags: ZKP_BATCHED / ZKP_SHPLONK / ZKP_NO_OPEN for every proof;
 * ZKS_CROSS_BATCH (with ZKP_BATCHED) merges the commitment MSMs of proofs in
 * flight into shared zk_msm_batch passes.  zks_batch_stats: merged passes,
 * passes holding more than one proof's groups, MSM items.  No reference
 * counterpart (src/lib.rs proves one circuit at a time). */
#define ZKS_EAGER 0x100
#define ZKS_CROSS_BATCH 0x200
typedef struct {
    int shape, err, z_ok, quotient_ok, ncommit, reserved;
    double queue_ms, run_ms, latency_ms, prove_ms;
} zks_result;
hipError_t zks_create(const zkp_circuit *shapes, int nshapes, int slots, size_t hbm_budget, int flags,
                      unsigned long long key_seed, void **srv);
int zks_submit(void *srv, int shape, unsigned long long seed);
hipError_t zks_wait(void *srv, int id, void *commitments, zks_result *res);
hipError_t zks_stats(void *srv, int *peak_running, size_t *peak_bytes, size_t *budget);
hipError_t zks_batch_stats(void *srv, unsigned long long *launches, unsigned long long *merged,
                           unsigned long long *items);
hipError_t zks_destroy(void *srv);

/* ---- host residency: pinned pool and device-to-host copy queue --------- */
/* Process-wide pinned host pool (replaces src/hugetlb.rs:17-229, the pinned
 * buffer cache behind HugePageAllocator): slabs of 2^28 bytes from
 * hipHostMalloc, each split into two 2^27-byte chunks (one Fr column of 2^22
 * rows).  A request of at most a chunk takes a chunk (partly used slabs
 * first), at most a slab a whole slab; larger requests, and frees of
 * pointers the pool did not hand out, return hipErrorInvalidValue.  Slabs
 * stay with the pool until the process ends; reserve pre-grows it
 * (reserve_pinned_buffer, hugetlb.rs:149). */
hipError_t zk_pinned_reserve(size_t slabs);
/* prepare_advice_buffer (src/lib.rs:107-112, src/buffer.rs:19-34): `count`
 * columns of `rows` Fr each (the circuit's advice columns, 2^k rows), zeroed,
 * from the pinned pool (so rows * 32 <= 2^28); cols (host array) receives the
 * pointers, which zkp_set_host_witness takes as the advice.  On failure no
 * column stays allocated.  zk_release_advice_buffer gives them back. */
hipError_t zk_prepare_advice_buffer(int rows, int count, void **cols);
hipError_t zk_release_advice_buffer(int count, void **cols);
hipError_t zk_pinned_alloc(size_t bytes, void **out);
hipError_t zk_pinned_free(void *p);
hipError_t zk_pinned_stats(size_t *slabs, size_t *chunks_in_use, size_t *slab_bytes, size_t *chunk_bytes);

/* Asynchronous device-to-host copy queue (replaces src/async_copy_queue.rs:8-90,
 * AsyncCopyQueue::new(28) at src/lib.rs:586).  A push is ordered after the
 * work enqueued so far on `producer` (an event, no host wait) and runs on the
 * queue's own stream; with max_inflight copies outstanding a push first waits
 * for the oldest.  flags: 0 = the caller leaves dev_src untouched until
 * sync_all; ZK_COPYQ_SNAPSHOT = copied first into the queue's staging buffer,
 * and `producer` waits for that device copy only (push_with_copy_buffer,
 * :42-55); ZK_COPYQ_RELEASE = the queue owns dev_src (hipMalloc'd) and frees
 * it after the copy (push / push_with_new_buffer, :57-82).  The two flags are
 * exclusive.  host_dst from zk_pinned_alloc makes the copy asynchronous. */
#define ZK_COPYQ_SNAPSHOT 1
#define ZK_COPYQ_RELEASE 2
hipError_t zk_copyq_create(int max_inflight, void **q);
hipError_t zk_copyq_push(void *q, void *host_dst, const void *dev_src, size_t bytes, hipStream_t producer,
                         int flags);
hipError_t zk_copyq_sync_all(void *q);
hipError_t zk_copyq_stats(void *q, unsigned long long *pushed, unsigned long long *full_waits,
                          double *max_wait_ms);
hipError_t zk_copyq_destroy(void *q);

#ifdef __cplusplus
}
#endif
#endif
