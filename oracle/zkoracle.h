/*
 * zkoracle -- CPU restatement of the zkWasm-prover halo2 KZG arithmetic core.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * path is compared against (tests/, the smoke check and the CPU-baseline leg of
 * bench.py).  It is never linked into, loaded by, or called from the product
 * library, and the product path has no CPU fallback to it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the reference repository root).  Elements are stored
 * exactly as halo2's `Fr([u64;4])` / `Fq([u64;4])`: four little-endian u64
 * limbs in Montgomery form with R = 2^256.  G1 affine = {x, y} (64 B, the
 * identity is (0, 0)); G1 Jacobian = {x, y, z} (96 B, the identity has z = 0).
 *
 * Pinning (what fixes each function independently of this file) is recorded in
 * the header comment of each test module; functions not pinned by anything
 * beyond a restatement are marked "parity unpinned" next to their declaration.
 */
#ifndef ZKORACLE_H
#define ZKORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { uint64_t l[4]; } zo_fr;
typedef struct { uint64_t l[4]; } zo_fp;
typedef struct { zo_fp x, y; } zo_g1a;
typedef struct { zo_fp x, y, z; } zo_g1j;

/* ---- threading ---------------------------------------------------------- */
void zo_set_threads(int n);
int zo_get_threads(void);

/* ---- Fr / Fp field (cuda/zprize/ff_dispatch_st.cuh:192-217, :428-529, :655-686) */
void zo_fr_add(zo_fr *r, const zo_fr *a, const zo_fr *b);
void zo_fr_sub(zo_fr *r, const zo_fr *a, const zo_fr *b);
void zo_fr_mul(zo_fr *r, const zo_fr *a, const zo_fr *b);
void zo_fr_inv(zo_fr *r, const zo_fr *a);          /* inv(0) = 0 */
void zo_fr_to_mont(zo_fr *r, const zo_fr *canonical);
void zo_fr_from_mont(zo_fr *r, const zo_fr *mont);
void zo_fr_pow(zo_fr *r, const zo_fr *a, uint64_t e);
void zo_fr_from_u64(zo_fr *r, uint64_t v);          /* Montgomery form of v */
void zo_fp_add(zo_fp *r, const zo_fp *a, const zo_fp *b);
void zo_fp_sub(zo_fp *r, const zo_fp *a, const zo_fp *b);
void zo_fp_mul(zo_fp *r, const zo_fp *a, const zo_fp *b);
void zo_fp_inv(zo_fp *r, const zo_fp *a);
void zo_fp_to_mont(zo_fp *r, const zo_fp *canonical);
void zo_fp_from_mont(zo_fp *r, const zo_fp *mont);

/* element-wise helpers over arrays (used by tests to build vectors) */
void zo_fr_to_mont_n(zo_fr *out, const zo_fr *in, size_t n);
void zo_fr_from_mont_n(zo_fr *out, const zo_fr *in, size_t n);

/* ---- seeded input generation ------------------------------------------- */
/* SplitMix64 stream -> 4 x u64 -> top limb masked to 254 bits -> reject >= r
 * -> Montgomery form.  Stream position advances per accepted draw. */
void zo_rand_fr(zo_fr *out, size_t n, uint64_t seed);
/* uniform nonzero variant (rejects 0 as well) */
void zo_rand_fr_nonzero(zo_fr *out, size_t n, uint64_t seed);
/* bases k_i * G with k_i = zo_rand_fr(seed) (exact scalar multiples) */
void zo_gen_bases_random(zo_g1a *out, size_t n, uint64_t seed);
/* bases (k0 + i*d) * G with k0, d drawn from seed: cheap for large n */
void zo_gen_bases_ap(zo_g1a *out, size_t n, uint64_t seed);

/* ---- G1 (cuda/zprize/ec.cuh:287-411) ------------------------------------ */
void zo_g1_generator(zo_g1a *g);
void zo_g1_dbl(zo_g1j *r, const zo_g1j *p);
void zo_g1_add(zo_g1j *r, const zo_g1j *p, const zo_g1j *q);
void zo_g1_add_mixed(zo_g1j *r, const zo_g1j *p, const zo_g1a *q);
void zo_g1_to_affine(zo_g1a *r, const zo_g1j *p);
void zo_g1_from_affine(zo_g1j *r, const zo_g1a *p);
void zo_g1_mul(zo_g1j *r, const zo_g1a *p, const zo_fr *s_mont);
int zo_g1a_is_on_curve(const zo_g1a *p);
/* batch Jacobian -> affine (one inversion per chunk) */
void zo_g1_batch_to_affine(zo_g1a *out, const zo_g1j *in, size_t n);

/* ---- MSM (cuda/msm.cu:9-467, src/cuda/msm.rs:239-455) ------------------ */
void zo_fr_inner_cycled(zo_fr *out, const zo_fr *s, size_t n, const zo_fr *c, size_t m);
void zo_msm_naive(zo_g1j *r, const zo_g1a *p, const zo_fr *s, size_t n);
/* Pippenger with unsigned c-bit windows over canonical scalars; c = 0 picks
 * a window from n.  Threads split the windows. */
void zo_msm(zo_g1j *r, const zo_g1a *p, const zo_fr *s, size_t n, int c);

/* ---- NTT (cuda/bn254.cu:1357-1428, src/cuda/ntt.rs:15-93) --------------- */
/* y_j = sum_i x_i omega^(i j), natural order in and out (halo2 best_fft).
 * n_inv != NULL scales every output by *n_inv (the iNTT divisor). */
void zo_ntt(zo_fr *a, size_t log_n, const zo_fr *omega, const zo_fr *n_inv);
/* O(n^2) definition, for small n only */
void zo_dft_naive(zo_fr *out, const zo_fr *in, size_t n, const zo_fr *omega);
/* omega for 2^log_n from the 2^28-th root 7^((r-1)/2^28) (halo2 EvaluationDomain) */
void zo_root_of_unity(zo_fr *omega, int log_n);
/* expand_omega_buffer natural layout: buf[i] = omega^i, i < n (cuda/bn254.cu:1073) */
void zo_expand_omega(zo_fr *buf, size_t n, const zo_fr *omega);

/* extended_prepare (cuda/bn254.cu:641-662, :1232-1253) */
void zo_extended_prepare(zo_fr *s, const zo_fr *coset_powers, int coset_powers_n,
                         size_t size, size_t extended_size, int to_coset);

/* ---- batch inversion / logup (cuda/bn254.cu:19-220, :1914-1993) -------- */
void zo_batch_invert(zo_fr *a, size_t n);            /* zeros stay zero */
void zo_logup_sum_input_inv(zo_fr *sum, const zo_fr *input, const zo_fr *beta,
                            int init, size_t n);
void zo_eval_logup_z_compose(zo_fr *z, const zo_fr *input_sum, const zo_fr *table,
                             const zo_fr *multiplicity, const zo_fr *beta,
                             zo_fr *last_z, int last_z_index, size_t n);
void zo_eval_logup_z_pure(zo_fr *z, zo_fr *last_z, int last_z_index, size_t n);

/* ---- grand products (src/permutation.rs:16-150, src/shuffle.rs:142-212) */
/* One permutation set: columns v[j], sigma[j] for j < ncols; delta_start =
 * DELTA^(set*chunk_len); writes z[0..n) as the exclusive running product of
 *   prod_j (v_j + beta*delta_start*DELTA^j*omega^i + gamma) /
 *   prod_j (v_j + beta*sigma_j + gamma);
 * rows > u are left as computed (the caller overwrites them with blinding). */
void zo_permutation_z_set(zo_fr *z, const zo_fr *const *v, const zo_fr *const *sigma,
                          int ncols, const zo_fr *beta, const zo_fr *gamma,
                          const zo_fr *delta_start, const zo_fr *delta,
                          const zo_fr *omega, size_t n);
/* Chains sets (src/permutation.rs:96-147): z_s[0..=u] *= prod_{s'<s} z_{s'}[u]
 * where z_{s'}[u] is taken before chaining. */
void zo_permutation_chain(zo_fr **zs, int nsets, size_t u);
/* shuffle group: ratio_i = prod_j (input_j+beta^(j+1)) / prod_j (shuffle_j+beta^(j+1));
 * z = exclusive running product (z[0] = 1). */
void zo_shuffle_z(zo_fr *z, const zo_fr *const *inputs, const zo_fr *const *shuffles,
                  int group_len, const zo_fr *beta, size_t n);

/* ---- eval_h per-row formulas (cuda/bn254.cu:542-1213) ------------------ */
void zo_field_op(zo_fr *res, const zo_fr *l, int l_rot, const zo_fr *l_c,
                 const zo_fr *r, int r_rot, const zo_fr *r_c, size_t n, int op);
/* terms: for term t, coeff[t] times prod over cols col[off[t]..off[t+1]) of
 * v[col][(i+rot)&(n-1)]; res[i] += sum_t */
void zo_field_op_batch_mul_sum(zo_fr *res, const zo_fr *const *vals, const int *rots,
                               const zo_fr *coeffs, const int *term_off, int nterms,
                               size_t n);
void zo_field_mul_zip(zo_fr *buf, const zo_fr *coeff, size_t coeff_n, size_t n);
void zo_permutation_eval_h_p1(zo_fr *res, const zo_fr *first, const zo_fr *last,
                              const zo_fr *l0, const zo_fr *l_last, const zo_fr *y, size_t n);
void zo_permutation_eval_h_p2(zo_fr *res, const zo_fr *const *set, const zo_fr *l0,
                              const zo_fr *y, int n_set, int rot, size_t n);
void zo_permutation_eval_h_l(zo_fr *res, const zo_fr *beta, const zo_fr *gamma,
                             const zo_fr *p, size_t n);
void zo_logup_eval_h(zo_fr *res, const zo_fr *input_product, const zo_fr *input_product_sum,
                     const zo_fr *table, const zo_fr *multiplicity, const zo_fr *z_first,
                     const zo_fr *z_last, const zo_fr *l0, const zo_fr *l_last,
                     const zo_fr *l_active_row, const zo_fr *y, int rot, size_t n);
void zo_logup_eval_h_extra_inputs(zo_fr *res, const zo_fr *input_product,
                                  const zo_fr *input_product_sum, const zo_fr *z,
                                  const zo_fr *l_active_row, const zo_fr *y, int rot, size_t n);
void zo_field_sum(zo_fr *res, const zo_fr *const *v, const zo_fr *const *v_c, const int *v_rot,
                  const zo_fr *omegas, int v_n, size_t n);
void zo_lookup_eval_h(zo_fr *res, const zo_fr *input, const zo_fr *table, const zo_fr *permuted_input,
                      const zo_fr *permuted_table, const zo_fr *z, const zo_fr *l0, const zo_fr *l_last,
                      const zo_fr *l_active_row, const zo_fr *y, const zo_fr *beta, const zo_fr *gamma,
                      int rot, size_t n);
void zo_logup_eval_h_inputs_product_sum(zo_fr *product, zo_fr *product_sum,
                                        const zo_fr *const *set, int n_set, size_t n);
void zo_shuffle_eval_h(zo_fr *res, const zo_fr *input, const zo_fr *table, const zo_fr *z,
                       const zo_fr *l0, const zo_fr *l_last, const zo_fr *l_active_row,
                       const zo_fr *y, int rot, size_t n);
void zo_shuffle_eval_h_v2(zo_fr *res, const zo_fr *const *inputs, const zo_fr *const *tables,
                          const zo_fr *betas, int group_len, const zo_fr *z, const zo_fr *l0,
                          const zo_fr *l_last, const zo_fr *l_active_row, const zo_fr *y,
                          int rot, size_t n);

/* multi-open */
void zo_poly_eval(zo_fr *res, const zo_fr *p, const zo_fr *x_table, size_t n);
void zo_kate_divide(zo_fr *poly, const zo_fr *z, size_t n);
void zo_shplonk_h_x_merge(zo_fr *res, const zo_fr *v, const zo_fr *values, const zo_fr *omegas,
                          const zo_fr *diff, int diff_n, size_t n);
void zo_shplonk_h_x_div_points(zo_fr *values, const zo_fr *omegas, const zo_fr *points,
                               int points_n, size_t n);

#ifdef __cplusplus
}
#endif
#endif
