/*
 * zkoracle.c -- CPU restatement of the zkWasm-prover arithmetic core.
 * TEST INFRASTRUCTURE ONLY (the checker; see zkoracle.h).
 */
#include "zkoracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------ */
/* Moduli.  BN254 scalar field r and base field p, little-endian u64 limbs. */
static const uint64_t R_MOD[4] = {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL,
                                  0xb85045b68181585dULL, 0x30644e72e131a029ULL};
static const uint64_t R_INV = 0xc2e1f593efffffffULL;   /* -r^-1 mod 2^64 */
static const uint64_t R_R2[4] = {0x1bb8e645ae216da7ULL, 0x53fe3ab1e35c59e3ULL,
                                 0x8c49833d53bb8085ULL, 0x0216d0b17f4e44a5ULL};
static const uint64_t P_MOD[4] = {0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL,
                                  0xb85045b68181585dULL, 0x30644e72e131a029ULL};
static const uint64_t P_INV = 0x87d20782e4866389ULL;   /* -p^-1 mod 2^64 */
static const uint64_t P_R2[4] = {0xf32cfc5b538afa89ULL, 0xb5e71911d44501fbULL,
                                 0x47ab1eff0a417ff6ULL, 0x06d89f71cab8351fULL};

static int g_threads = 0;
void zo_set_threads(int n) { g_threads = n; }
int zo_get_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}
#define OMP_THREADS num_threads(zo_get_threads())

/* ---- generic 4-limb Montgomery arithmetic ------------------------------- */
static inline int geq4(const uint64_t a[4], const uint64_t m[4]) {
    for (int i = 3; i >= 0; i--) {
        if (a[i] > m[i]) return 1;
        if (a[i] < m[i]) return 0;
    }
    return 1;
}
static inline void sub4(uint64_t r[4], const uint64_t a[4], const uint64_t b[4]) {
    u128 borrow = 0;
    for (int i = 0; i < 4; i++) {
        u128 t = (u128)a[i] - b[i] - (uint64_t)borrow;
        r[i] = (uint64_t)t;
        borrow = (t >> 64) ? 1 : 0;
    }
}
static inline void madd(uint64_t r[4], const uint64_t a[4], const uint64_t b[4],
                        const uint64_t m[4]) {
    u128 c = 0;
    uint64_t t[4];
    for (int i = 0; i < 4; i++) {
        c += (u128)a[i] + b[i];
        t[i] = (uint64_t)c;
        c >>= 64;
    }
    if (geq4(t, m)) sub4(t, t, m);
    memcpy(r, t, 32);
}
static inline void msub(uint64_t r[4], const uint64_t a[4], const uint64_t b[4],
                        const uint64_t m[4]) {
    uint64_t t[4];
    u128 borrow = 0;
    for (int i = 0; i < 4; i++) {
        u128 x = (u128)a[i] - b[i] - (uint64_t)borrow;
        t[i] = (uint64_t)x;
        borrow = (x >> 64) ? 1 : 0;
    }
    if (borrow) {
        u128 c = 0;
        for (int i = 0; i < 4; i++) {
            c += (u128)t[i] + m[i];
            t[i] = (uint64_t)c;
            c >>= 64;
        }
    }
    memcpy(r, t, 32);
}
/* CIOS Montgomery product a*b*2^-256 mod m; inputs < m, output < m. */
static inline void mmul(uint64_t r[4], const uint64_t a[4], const uint64_t b[4],
                        const uint64_t m[4], uint64_t inv) {
    uint64_t t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; i++) {
        u128 c = 0;
        for (int j = 0; j < 4; j++) {
            c = (u128)a[j] * b[i] + t[j] + (uint64_t)(c >> 64);
            t[j] = (uint64_t)c;
        }
        c = (u128)t[4] + (uint64_t)(c >> 64);
        t[4] = (uint64_t)c;
        t[5] = (uint64_t)(c >> 64);
        uint64_t q = t[0] * inv;
        c = (u128)q * m[0] + t[0];
        for (int j = 1; j < 4; j++) {
            c = (u128)q * m[j] + t[j] + (uint64_t)(c >> 64);
            t[j - 1] = (uint64_t)c;
        }
        c = (u128)t[4] + (uint64_t)(c >> 64);
        t[3] = (uint64_t)c;
        t[4] = t[5] + (uint64_t)(c >> 64);
    }
    if (t[4] || geq4(t, m)) sub4(t, t, m);
    memcpy(r, t, 32);
}
static inline int is_zero4(const uint64_t a[4]) { return !(a[0] | a[1] | a[2] | a[3]); }

/* ---- Fr ---------------------------------------------------------------- */
static inline void fr_add(zo_fr *r, const zo_fr *a, const zo_fr *b) { madd(r->l, a->l, b->l, R_MOD); }
static inline void fr_sub(zo_fr *r, const zo_fr *a, const zo_fr *b) { msub(r->l, a->l, b->l, R_MOD); }
static inline void fr_mul(zo_fr *r, const zo_fr *a, const zo_fr *b) { mmul(r->l, a->l, b->l, R_MOD, R_INV); }
static void fr_one(zo_fr *r) { static const zo_fr c1 = {{1, 0, 0, 0}}; mmul(r->l, c1.l, R_R2, R_MOD, R_INV); }
static void fr_pow_big(zo_fr *r, const zo_fr *a, const uint64_t e[4]) {
    zo_fr acc; fr_one(&acc);
    zo_fr base = *a;
    for (int i = 0; i < 4; i++)
        for (int b = 0; b < 64; b++) {
            if ((e[i] >> b) & 1) fr_mul(&acc, &acc, &base);
            fr_mul(&base, &base, &base);
        }
    *r = acc;
}
void zo_fr_add(zo_fr *r, const zo_fr *a, const zo_fr *b) { fr_add(r, a, b); }
void zo_fr_sub(zo_fr *r, const zo_fr *a, const zo_fr *b) { fr_sub(r, a, b); }
void zo_fr_mul(zo_fr *r, const zo_fr *a, const zo_fr *b) { fr_mul(r, a, b); }
void zo_fr_inv(zo_fr *r, const zo_fr *a) {
    /* Fermat: a^(r-2); 0 maps to 0 like the device inverse (ff_dispatch_st.cuh:655) */
    uint64_t e[4]; memcpy(e, R_MOD, 32); e[0] -= 2;
    fr_pow_big(r, a, e);
}
void zo_fr_to_mont(zo_fr *r, const zo_fr *c) { mmul(r->l, c->l, R_R2, R_MOD, R_INV); }
void zo_fr_from_mont(zo_fr *r, const zo_fr *m) { static const uint64_t one[4] = {1, 0, 0, 0}; mmul(r->l, m->l, one, R_MOD, R_INV); }
void zo_fr_pow(zo_fr *r, const zo_fr *a, uint64_t e) { uint64_t ee[4] = {e, 0, 0, 0}; fr_pow_big(r, a, ee); }
void zo_fr_from_u64(zo_fr *r, uint64_t v) { zo_fr c = {{v, 0, 0, 0}}; zo_fr_to_mont(r, &c); }
void zo_fr_to_mont_n(zo_fr *out, const zo_fr *in, size_t n) {
#pragma omp parallel for OMP_THREADS if (n > 4096)
    for (size_t i = 0; i < n; i++) zo_fr_to_mont(&out[i], &in[i]);
}
void zo_fr_from_mont_n(zo_fr *out, const zo_fr *in, size_t n) {
#pragma omp parallel for OMP_THREADS if (n > 4096)
    for (size_t i = 0; i < n; i++) zo_fr_from_mont(&out[i], &in[i]);
}

/* ---- Fp ---------------------------------------------------------------- */
static inline void fp_add(zo_fp *r, const zo_fp *a, const zo_fp *b) { madd(r->l, a->l, b->l, P_MOD); }
static inline void fp_sub(zo_fp *r, const zo_fp *a, const zo_fp *b) { msub(r->l, a->l, b->l, P_MOD); }
static inline void fp_mul(zo_fp *r, const zo_fp *a, const zo_fp *b) { mmul(r->l, a->l, b->l, P_MOD, P_INV); }
static inline void fp_sqr(zo_fp *r, const zo_fp *a) { mmul(r->l, a->l, a->l, P_MOD, P_INV); }
static inline void fp_dbl(zo_fp *r, const zo_fp *a) { madd(r->l, a->l, a->l, P_MOD); }
static inline int fp_is_zero(const zo_fp *a) { return is_zero4(a->l); }
static inline int fp_eq(const zo_fp *a, const zo_fp *b) { return memcmp(a, b, 32) == 0; }
static void fp_one(zo_fp *r) { static const uint64_t c1[4] = {1, 0, 0, 0}; mmul(r->l, c1, P_R2, P_MOD, P_INV); }
void zo_fp_add(zo_fp *r, const zo_fp *a, const zo_fp *b) { fp_add(r, a, b); }
void zo_fp_sub(zo_fp *r, const zo_fp *a, const zo_fp *b) { fp_sub(r, a, b); }
void zo_fp_mul(zo_fp *r, const zo_fp *a, const zo_fp *b) { fp_mul(r, a, b); }
void zo_fp_inv(zo_fp *r, const zo_fp *a) {
    uint64_t e[4]; memcpy(e, P_MOD, 32); e[0] -= 2;
    zo_fp acc; fp_one(&acc);
    zo_fp base = *a;
    for (int i = 0; i < 4; i++)
        for (int b = 0; b < 64; b++) {
            if ((e[i] >> b) & 1) fp_mul(&acc, &acc, &base);
            fp_sqr(&base, &base);
        }
    *r = acc;
}
void zo_fp_to_mont(zo_fp *r, const zo_fp *c) { mmul(r->l, c->l, P_R2, P_MOD, P_INV); }
void zo_fp_from_mont(zo_fp *r, const zo_fp *m) { static const uint64_t one[4] = {1, 0, 0, 0}; mmul(r->l, m->l, one, P_MOD, P_INV); }

/* ---- RNG --------------------------------------------------------------- */
static inline uint64_t splitmix64(uint64_t *s) {
    uint64_t z = (*s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
static void rand_fr_impl(zo_fr *out, size_t n, uint64_t seed, int nonzero) {
    uint64_t s = seed;
    for (size_t i = 0; i < n;) {
        zo_fr c;
        for (int j = 0; j < 4; j++) c.l[j] = splitmix64(&s);
        c.l[3] &= (1ULL << 62) - 1;
        if (geq4(c.l, R_MOD)) continue;
        if (nonzero && is_zero4(c.l)) continue;
        zo_fr_to_mont(&out[i], &c);
        i++;
    }
}
void zo_rand_fr(zo_fr *out, size_t n, uint64_t seed) { rand_fr_impl(out, n, seed, 0); }
void zo_rand_fr_nonzero(zo_fr *out, size_t n, uint64_t seed) { rand_fr_impl(out, n, seed, 1); }

/* ---- G1 (y^2 = x^3 + 3 over Fp) ----------------------------------------- */
void zo_g1_generator(zo_g1a *g) {
    zo_fp one = {{1, 0, 0, 0}}, two = {{2, 0, 0, 0}};
    zo_fp_to_mont(&g->x, &one);
    zo_fp_to_mont(&g->y, &two);
}
static inline int g1a_is_id(const zo_g1a *p) { return fp_is_zero(&p->x) && fp_is_zero(&p->y); }
static inline void g1j_set_id(zo_g1j *r) { memset(r, 0, sizeof(*r)); fp_one(&r->y); }

/* dbl-2009-l (a = 0): the formula cuda/zprize/ec.cuh:287-314 uses */
void zo_g1_dbl(zo_g1j *r, const zo_g1j *p) {
    if (fp_is_zero(&p->z)) { *r = *p; return; }
    zo_fp A, B, C, D, E, F, t;
    fp_sqr(&A, &p->x);
    fp_sqr(&B, &p->y);
    fp_sqr(&C, &B);
    fp_add(&t, &p->x, &B); fp_sqr(&t, &t); fp_sub(&t, &t, &A); fp_sub(&t, &t, &C); fp_dbl(&D, &t);
    fp_dbl(&E, &A); fp_add(&E, &E, &A);
    fp_sqr(&F, &E);
    zo_fp X3, Y3, Z3;
    fp_dbl(&t, &D); fp_sub(&X3, &F, &t);
    fp_sub(&t, &D, &X3); fp_mul(&Y3, &E, &t);
    fp_dbl(&t, &C); fp_dbl(&t, &t); fp_dbl(&t, &t); fp_sub(&Y3, &Y3, &t);
    fp_mul(&Z3, &p->y, &p->z); fp_dbl(&Z3, &Z3);
    r->x = X3; r->y = Y3; r->z = Z3;
}
/* add-2007-bl with the zero / equal-point cases (ec.cuh:322-369) */
void zo_g1_add(zo_g1j *r, const zo_g1j *p, const zo_g1j *q) {
    if (fp_is_zero(&p->z)) { *r = *q; return; }
    if (fp_is_zero(&q->z)) { *r = *p; return; }
    zo_fp Z1Z1, Z2Z2, U1, U2, S1, S2, H, I, J, rr, V, t;
    fp_sqr(&Z1Z1, &p->z);
    fp_sqr(&Z2Z2, &q->z);
    fp_mul(&U1, &p->x, &Z2Z2);
    fp_mul(&U2, &q->x, &Z1Z1);
    fp_mul(&S1, &p->y, &q->z); fp_mul(&S1, &S1, &Z2Z2);
    fp_mul(&S2, &q->y, &p->z); fp_mul(&S2, &S2, &Z1Z1);
    if (fp_eq(&U1, &U2)) {
        if (fp_eq(&S1, &S2)) { zo_g1_dbl(r, p); return; }
        g1j_set_id(r); return;
    }
    fp_sub(&H, &U2, &U1);
    fp_dbl(&I, &H); fp_sqr(&I, &I);
    fp_mul(&J, &H, &I);
    fp_sub(&rr, &S2, &S1); fp_dbl(&rr, &rr);
    fp_mul(&V, &U1, &I);
    zo_fp X3, Y3, Z3;
    fp_sqr(&X3, &rr); fp_sub(&X3, &X3, &J); fp_dbl(&t, &V); fp_sub(&X3, &X3, &t);
    fp_sub(&t, &V, &X3); fp_mul(&Y3, &rr, &t); fp_mul(&t, &S1, &J); fp_dbl(&t, &t); fp_sub(&Y3, &Y3, &t);
    fp_add(&t, &p->z, &q->z); fp_sqr(&t, &t); fp_sub(&t, &t, &Z1Z1); fp_sub(&t, &t, &Z2Z2); fp_mul(&Z3, &t, &H);
    r->x = X3; r->y = Y3; r->z = Z3;
}
/* madd-2007-bl (Jacobian + affine), the affine identity handled */
void zo_g1_add_mixed(zo_g1j *r, const zo_g1j *p, const zo_g1a *q) {
    if (g1a_is_id(q)) { *r = *p; return; }
    if (fp_is_zero(&p->z)) { r->x = q->x; r->y = q->y; fp_one(&r->z); return; }
    zo_fp Z1Z1, U2, S2, H, HH, I, J, rr, V, t;
    fp_sqr(&Z1Z1, &p->z);
    fp_mul(&U2, &q->x, &Z1Z1);
    fp_mul(&S2, &q->y, &p->z); fp_mul(&S2, &S2, &Z1Z1);
    if (fp_eq(&U2, &p->x)) {
        if (fp_eq(&S2, &p->y)) { zo_g1_dbl(r, p); return; }
        g1j_set_id(r); return;
    }
    fp_sub(&H, &U2, &p->x);
    fp_sqr(&HH, &H);
    fp_dbl(&I, &HH); fp_dbl(&I, &I);
    fp_mul(&J, &H, &I);
    fp_sub(&rr, &S2, &p->y); fp_dbl(&rr, &rr);
    fp_mul(&V, &p->x, &I);
    zo_fp X3, Y3, Z3;
    fp_sqr(&X3, &rr); fp_sub(&X3, &X3, &J); fp_dbl(&t, &V); fp_sub(&X3, &X3, &t);
    fp_sub(&t, &V, &X3); fp_mul(&Y3, &rr, &t); fp_mul(&t, &p->y, &J); fp_dbl(&t, &t); fp_sub(&Y3, &Y3, &t);
    fp_add(&t, &p->z, &H); fp_sqr(&t, &t); fp_sub(&t, &t, &Z1Z1); fp_sub(&Z3, &t, &HH);
    r->x = X3; r->y = Y3; r->z = Z3;
}
void zo_g1_to_affine(zo_g1a *r, const zo_g1j *p) {
    if (fp_is_zero(&p->z)) { memset(r, 0, sizeof(*r)); return; }
    zo_fp zi, zi2, zi3;
    zo_fp_inv(&zi, &p->z);
    fp_sqr(&zi2, &zi);
    fp_mul(&zi3, &zi2, &zi);
    fp_mul(&r->x, &p->x, &zi2);
    fp_mul(&r->y, &p->y, &zi3);
}
void zo_g1_from_affine(zo_g1j *r, const zo_g1a *p) {
    if (g1a_is_id(p)) { g1j_set_id(r); return; }
    r->x = p->x; r->y = p->y; fp_one(&r->z);
}
void zo_g1_mul(zo_g1j *r, const zo_g1a *p, const zo_fr *s_mont) {
    zo_fr s; zo_fr_from_mont(&s, s_mont);
    zo_g1j acc; g1j_set_id(&acc);
    for (int i = 255; i >= 0; i--) {
        zo_g1_dbl(&acc, &acc);
        if ((s.l[i / 64] >> (i % 64)) & 1) zo_g1_add_mixed(&acc, &acc, p);
    }
    *r = acc;
}
int zo_g1a_is_on_curve(const zo_g1a *p) {
    if (g1a_is_id(p)) return 1;
    zo_fp y2, x3, b = {{3, 0, 0, 0}};
    zo_fp_to_mont(&b, &b);
    fp_sqr(&y2, &p->y);
    fp_sqr(&x3, &p->x); fp_mul(&x3, &x3, &p->x); fp_add(&x3, &x3, &b);
    return fp_eq(&y2, &x3);
}
void zo_g1_batch_to_affine(zo_g1a *out, const zo_g1j *in, size_t n) {
    const size_t chunk = 4096;
    size_t nchunks = (n + chunk - 1) / chunk;
#pragma omp parallel for OMP_THREADS schedule(dynamic)
    for (size_t c = 0; c < nchunks; c++) {
        size_t s = c * chunk, e = s + chunk < n ? s + chunk : n;
        zo_fp *pre = (zo_fp *)malloc(sizeof(zo_fp) * (e - s));
        zo_fp acc; fp_one(&acc);
        for (size_t i = s; i < e; i++) {
            pre[i - s] = acc;
            if (!fp_is_zero(&in[i].z)) fp_mul(&acc, &acc, &in[i].z);
        }
        zo_fp inv; zo_fp_inv(&inv, &acc);
        for (size_t i = e; i-- > s;) {
            if (fp_is_zero(&in[i].z)) { memset(&out[i], 0, sizeof(zo_g1a)); continue; }
            zo_fp zi, zi2, zi3;
            fp_mul(&zi, &inv, &pre[i - s]);
            fp_mul(&inv, &inv, &in[i].z);
            fp_sqr(&zi2, &zi); fp_mul(&zi3, &zi2, &zi);
            fp_mul(&out[i].x, &in[i].x, &zi2);
            fp_mul(&out[i].y, &in[i].y, &zi3);
        }
        free(pre);
    }
}
void zo_gen_bases_random(zo_g1a *out, size_t n, uint64_t seed) {
    zo_fr *k = (zo_fr *)malloc(sizeof(zo_fr) * n);
    zo_g1j *j = (zo_g1j *)malloc(sizeof(zo_g1j) * n);
    zo_rand_fr(k, n, seed);
    zo_g1a g; zo_g1_generator(&g);
#pragma omp parallel for OMP_THREADS schedule(dynamic, 64)
    for (size_t i = 0; i < n; i++) zo_g1_mul(&j[i], &g, &k[i]);
    zo_g1_batch_to_affine(out, j, n);
    free(k); free(j);
}
void zo_gen_bases_ap(zo_g1a *out, size_t n, uint64_t seed) {
    zo_fr kd[2];
    zo_rand_fr(kd, 2, seed);
    zo_g1a g; zo_g1_generator(&g);
    zo_g1j dj; zo_g1_mul(&dj, &g, &kd[1]);
    zo_g1a da; zo_g1_to_affine(&da, &dj);
    zo_g1j *j = (zo_g1j *)malloc(sizeof(zo_g1j) * n);
    const size_t chunk = 1 << 14;
    size_t nchunks = (n + chunk - 1) / chunk;
#pragma omp parallel for OMP_THREADS schedule(dynamic)
    for (size_t c = 0; c < nchunks; c++) {
        /* start = (k0 + s*d) G */
        size_t s = c * chunk, e = s + chunk < n ? s + chunk : n;
        zo_fr si, ks; zo_fr_from_u64(&si, (uint64_t)s);
        fr_mul(&ks, &si, &kd[1]); fr_add(&ks, &ks, &kd[0]);
        zo_g1j acc; zo_g1_mul(&acc, &g, &ks);
        for (size_t i = s; i < e; i++) {
            j[i] = acc;
            zo_g1_add_mixed(&acc, &acc, &da);
        }
    }
    zo_g1_batch_to_affine(out, j, n);
    free(j);
}

/* ---- MSM ----------------------------------------------------------------- */
/* out = sum_i s[i] * c[i mod m]: the closed-form expectation of the
 * reference's test_msm (src/cuda/msm.rs:477-511). */
void zo_fr_inner_cycled(zo_fr *out, const zo_fr *s, size_t n, const zo_fr *c, size_t m) {
    zo_fr *part = (zo_fr *)calloc(m, sizeof(zo_fr));
    for (size_t i = 0; i < n; i++) fr_add(&part[i % m], &part[i % m], &s[i]);
    zo_fr acc = {{0, 0, 0, 0}};
    for (size_t j = 0; j < m; j++) {
        zo_fr t; fr_mul(&t, &part[j], &c[j]); fr_add(&acc, &acc, &t);
    }
    *out = acc;
    free(part);
}
void zo_msm_naive(zo_g1j *r, const zo_g1a *p, const zo_fr *s, size_t n) {
    zo_g1j acc; g1j_set_id(&acc);
    for (size_t i = 0; i < n; i++) {
        zo_g1j t; zo_g1_mul(&t, &p[i], &s[i]);
        zo_g1_add(&acc, &acc, &t);
    }
    *r = acc;
}
static inline uint32_t get_bits(const zo_fr *c, unsigned start, unsigned bits) {
    unsigned limb = start / 64, off = start % 64;
    u128 w = c->l[limb];
    if (limb + 1 < 4) w |= (u128)c->l[limb + 1] << 64;
    return (uint32_t)((w >> off) & ((1ULL << bits) - 1));
}
/* Unsigned-window Pippenger: the bucket semantics of cuda/msm.cu:9-181
 * (digit d of window w adds P_i into bucket (w, d); d = 0 is skipped), then
 * sum_d d*B_d per window by running sums and Horner over windows
 * (cuda/msm.cu:246-318).  Parallel over (point range, window) tasks: each
 * task fills its own buckets from its range and reduces them to its share
 * S_w^(r) = sum_d d*B_{w,d}^(r); S_w = sum_r S_w^(r) (bucket sums are linear
 * in the points), so every host core has work even when the window count is
 * below the thread count.  The point is the same for any split. */
void zo_msm(zo_g1j *r, const zo_g1a *p, const zo_fr *s, size_t n, int c) {
    const int threads = zo_get_threads();
    size_t ranges = 1;
    if (c <= 0) {
        /* ranges so that there are >= 2 tasks per thread, then the window for
           one range's points */
        int c0 = 4;
        while (c0 < 16 && ((size_t)1 << (c0 + 3)) < n) c0++;
        const size_t w0 = (254 + c0 - 1) / c0;
        ranges = (2 * (size_t)threads + w0 - 1) / w0;
        if (ranges > n / 4096) ranges = n / 4096;
        if (ranges < 1) ranges = 1;
        const size_t nr = (n + ranges - 1) / ranges;
        c = 4;
        while (c < 16 && ((size_t)1 << (c + 3)) < nr) c++;
    }
    const unsigned bits = 254;
    const int windows = (bits + c - 1) / c;
    zo_fr *canon = (zo_fr *)malloc(sizeof(zo_fr) * (n ? n : 1));
    zo_fr_from_mont_n(canon, s, n);
    zo_g1j *part = (zo_g1j *)malloc(sizeof(zo_g1j) * windows * ranges);
    const size_t nb = (size_t)1 << c;
    const long tasks = (long)windows * (long)ranges;
#pragma omp parallel for OMP_THREADS schedule(dynamic, 1)
    for (long t = 0; t < tasks; t++) {
        const int w = (int)(t % windows);
        const size_t rg = (size_t)(t / windows);
        const size_t lo = n * rg / ranges, hi = n * (rg + 1) / ranges;
        zo_g1j *b = (zo_g1j *)malloc(sizeof(zo_g1j) * nb);
        for (size_t i = 0; i < nb; i++) g1j_set_id(&b[i]);
        unsigned start = (unsigned)w * c;
        unsigned nbits = bits - start < (unsigned)c ? bits - start : (unsigned)c;
        for (size_t i = lo; i < hi; i++) {
            uint32_t d = get_bits(&canon[i], start, nbits);
            if (d) zo_g1_add_mixed(&b[d], &b[d], &p[i]);
        }
        zo_g1j run, acc; g1j_set_id(&run); g1j_set_id(&acc);
        for (size_t d = nb - 1; d >= 1; d--) {
            zo_g1_add(&run, &run, &b[d]);
            zo_g1_add(&acc, &acc, &run);
        }
        part[t] = acc;
        free(b);
    }
    zo_g1j *wsum = (zo_g1j *)malloc(sizeof(zo_g1j) * windows);
    for (int w = 0; w < windows; w++) {
        wsum[w] = part[w];
        for (size_t rg = 1; rg < ranges; rg++) zo_g1_add(&wsum[w], &wsum[w], &part[rg * windows + w]);
    }
    zo_g1j acc = wsum[windows - 1];
    for (int w = windows - 2; w >= 0; w--) {
        for (int k = 0; k < c; k++) zo_g1_dbl(&acc, &acc);
        zo_g1_add(&acc, &acc, &wsum[w]);
    }
    *r = acc;
    free(wsum); free(part); free(canon);
}

/* ---- NTT ----------------------------------------------------------------- */
void zo_root_of_unity(zo_fr *omega, int log_n) {
    /* 7^((r-1)/2^28) is the 2^28-th root; square it (28 - log_n) times */
    static const zo_fr root28_c = {{0xd34f1ed960c37c9cULL, 0x3215cf6dd39329c8ULL,
                                    0x98865ea93dd31f74ULL, 0x03ddb9f5166d18b7ULL}};
    zo_fr w; zo_fr_to_mont(&w, &root28_c);
    for (int i = log_n; i < 28; i++) fr_mul(&w, &w, &w);
    *omega = w;
}
void zo_expand_omega(zo_fr *buf, size_t n, const zo_fr *omega) {
    zo_fr acc; fr_one(&acc);
    for (size_t i = 0; i < n; i++) { buf[i] = acc; fr_mul(&acc, &acc, omega); }
}
void zo_dft_naive(zo_fr *out, const zo_fr *in, size_t n, const zo_fr *omega) {
    zo_fr wj; fr_one(&wj);
    for (size_t j = 0; j < n; j++) {
        zo_fr acc = {{0, 0, 0, 0}}, x; fr_one(&x);
        for (size_t i = 0; i < n; i++) {
            zo_fr t; fr_mul(&t, &in[i], &x); fr_add(&acc, &acc, &t);
            fr_mul(&x, &x, &wj);
        }
        out[j] = acc;
        fr_mul(&wj, &wj, omega);
    }
}
static inline size_t bitrev(size_t x, int bits) {
    size_t r = 0;
    for (int i = 0; i < bits; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}
/* radix-2 DIT after a bit-reversal permutation: output is the DFT in natural
 * order, which is what `ntt` returns for natural-order input
 * (cuda/bn254.cu:1357-1428; halo2 best_fft semantics, src/profile.rs:220-247) */
void zo_ntt(zo_fr *a, size_t log_n, const zo_fr *omega, const zo_fr *n_inv) {
    size_t n = (size_t)1 << log_n;
    for (size_t i = 0; i < n; i++) {
        size_t j = bitrev(i, (int)log_n);
        if (j > i) { zo_fr t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    zo_fr *tw = (zo_fr *)malloc(sizeof(zo_fr) * (n / 2 + 1));
    zo_expand_omega(tw, n / 2 + 1, omega);
    int nt = zo_get_threads();
    for (size_t m = 1; m < n; m <<= 1) {
        size_t step = n / (2 * m);
        long long nblocks = (long long)(n / (2 * m));
        if (nblocks >= nt || m < 64) {
#pragma omp parallel for OMP_THREADS if (n >= 4096)
            for (long long bk = 0; bk < nblocks; bk++) {
                size_t k = (size_t)bk * 2 * m;
                for (size_t j = 0; j < m; j++) {
                    zo_fr t, u = a[k + j];
                    fr_mul(&t, &a[k + j + m], &tw[j * step]);
                    fr_add(&a[k + j], &u, &t);
                    fr_sub(&a[k + j + m], &u, &t);
                }
            }
        } else {
            for (long long bk = 0; bk < nblocks; bk++) {
                size_t k = (size_t)bk * 2 * m;
#pragma omp parallel for OMP_THREADS if (m >= 4096)
                for (long long jj = 0; jj < (long long)m; jj++) {
                    size_t j = (size_t)jj;
                    zo_fr t, u = a[k + j];
                    fr_mul(&t, &a[k + j + m], &tw[j * step]);
                    fr_add(&a[k + j], &u, &t);
                    fr_sub(&a[k + j + m], &u, &t);
                }
            }
        }
    }
    free(tw);
    if (n_inv) {
#pragma omp parallel for OMP_THREADS if (n >= 4096)
        for (size_t i = 0; i < n; i++) fr_mul(&a[i], &a[i], n_inv);
    }
}

/* cuda/bn254.cu:641-662 with the wrapper at :1232-1253: with to_coset == 0
 * zero-pad [size, extended_size) and scale [0, size); otherwise scale all of
 * [0, extended_size).  s[i] *= coset_powers[(i % coset_powers_n) - 1] when
 * i % coset_powers_n != 0. */
void zo_extended_prepare(zo_fr *s, const zo_fr *coset_powers, int coset_powers_n,
                         size_t size, size_t extended_size, int to_coset) {
    size_t lim = to_coset ? extended_size : size;
    if (!to_coset) memset(&s[size], 0, sizeof(zo_fr) * (extended_size - size));
#pragma omp parallel for OMP_THREADS if (lim >= 4096)
    for (size_t i = 0; i < lim; i++) {
        int idx = (int)(i % (size_t)coset_powers_n);
        if (idx != 0) fr_mul(&s[i], &s[i], &coset_powers[idx - 1]);
    }
}

/* ---- batch inversion / logup --------------------------------------------- */
/* Montgomery's trick over contiguous chunks, one chunk per task (halo2's
 * BatchInvert run per parallel chunk, as src/permutation.rs:73-75 and
 * src/shuffle.rs:163-165 do); zeros stay zero.  The inverses do not depend
 * on the split. */
void zo_batch_invert(zo_fr *a, size_t n) {
    zo_fr *pre = (zo_fr *)malloc(sizeof(zo_fr) * (n ? n : 1));
    const size_t chunks = n >= 65536 ? (size_t)zo_get_threads() * 4 : 1;
#pragma omp parallel for OMP_THREADS schedule(dynamic, 1) if (chunks > 1)
    for (long c = 0; c < (long)chunks; c++) {
        const size_t lo = n * (size_t)c / chunks, hi = n * (size_t)(c + 1) / chunks;
        zo_fr acc; fr_one(&acc);
        for (size_t i = lo; i < hi; i++) {
            pre[i] = acc;
            if (!is_zero4(a[i].l)) fr_mul(&acc, &acc, &a[i]);
        }
        zo_fr inv; zo_fr_inv(&inv, &acc);
        for (size_t i = hi; i-- > lo;) {
            if (is_zero4(a[i].l)) continue;
            zo_fr t; fr_mul(&t, &inv, &pre[i]);
            fr_mul(&inv, &inv, &a[i]);
            a[i] = t;
        }
    }
    free(pre);
}
/* cuda/bn254.cu:19-61 + :1914-1934: sum[i] (= or +=) 1 / (input[i] + beta) */
void zo_logup_sum_input_inv(zo_fr *sum, const zo_fr *input, const zo_fr *beta,
                            int init, size_t n) {
    zo_fr *t = (zo_fr *)malloc(sizeof(zo_fr) * n);
#pragma omp parallel for OMP_THREADS if (n >= 4096)
    for (size_t i = 0; i < n; i++) fr_add(&t[i], &input[i], beta);
    zo_batch_invert(t, n);
#pragma omp parallel for OMP_THREADS if (n >= 4096)
    for (size_t i = 0; i < n; i++) {
        if (init == 0) sum[i] = t[i];
        else fr_add(&sum[i], &sum[i], &t[i]);
    }
    free(t);
}
/* cuda/bn254.cu:119-202 + :1936-1966: z[i] = carry + sum_{j<i} z_in[j] with
 * carry = last_z[0]; afterwards last_z[0] = z[last_z_index]. */
void zo_eval_logup_z_pure(zo_fr *z, zo_fr *last_z, int last_z_index, size_t n) {
    zo_fr acc = last_z[0];
    for (size_t i = 0; i < n; i++) {
        zo_fr v = z[i];
        z[i] = acc;
        fr_add(&acc, &acc, &v);
    }
    last_z[0] = z[last_z_index];
}
/* cuda/bn254.cu:1968-1993: z[i] = input_sum[i] - m[i] / (table[i] + beta), then
 * the exclusive scan above. */
void zo_eval_logup_z_compose(zo_fr *z, const zo_fr *input_sum, const zo_fr *table,
                             const zo_fr *multiplicity, const zo_fr *beta,
                             zo_fr *last_z, int last_z_index, size_t n) {
    zo_fr *t = (zo_fr *)malloc(sizeof(zo_fr) * n);
    for (size_t i = 0; i < n; i++) fr_add(&t[i], &table[i], beta);
    zo_batch_invert(t, n);
    for (size_t i = 0; i < n; i++) {
        zo_fr u; fr_mul(&u, &t[i], &multiplicity[i]);
        fr_sub(&z[i], &input_sum[i], &u);
    }
    free(t);
    zo_eval_logup_z_pure(z, last_z, last_z_index, n);
}

/* ---- grand products ------------------------------------------------------ */
/* src/permutation.rs:44-128 for one chunk of columns */
void zo_permutation_z_set(zo_fr *z, const zo_fr *const *v, const zo_fr *const *sigma,
                          int ncols, const zo_fr *beta, const zo_fr *gamma,
                          const zo_fr *delta_start, const zo_fr *delta,
                          const zo_fr *omega, size_t n) {
    zo_fr *den = (zo_fr *)malloc(sizeof(zo_fr) * n);
    for (size_t i = 0; i < n; i++) {
        zo_fr acc; fr_one(&acc);
        for (int j = 0; j < ncols; j++) {
            zo_fr t; fr_mul(&t, beta, &sigma[j][i]); fr_add(&t, &t, gamma); fr_add(&t, &t, &v[j][i]);
            fr_mul(&acc, &acc, &t);
        }
        den[i] = acc;
    }
    zo_batch_invert(den, n);
    zo_fr dcol = *delta_start;
    for (int j = 0; j < ncols; j++) {
        zo_fr x; fr_mul(&x, &dcol, beta); /* beta * delta^(s*L+j) * omega^i */
        for (size_t i = 0; i < n; i++) {
            zo_fr t; fr_add(&t, &x, gamma); fr_add(&t, &t, &v[j][i]);
            fr_mul(&den[i], &den[i], &t);
            fr_mul(&x, &x, omega);
        }
        fr_mul(&dcol, &dcol, delta);
    }
    zo_fr acc; fr_one(&acc);
    for (size_t i = 0; i < n; i++) {
        z[i] = acc;
        fr_mul(&acc, &acc, &den[i]);
    }
    free(den);
}
void zo_permutation_chain(zo_fr **zs, int nsets, size_t u) {
    zo_fr carry; fr_one(&carry);
    for (int s = 0; s < nsets; s++) {
        zo_fr tail = zs[s][u];
        if (s > 0)
            for (size_t i = 0; i <= u; i++) fr_mul(&zs[s][i], &zs[s][i], &carry);
        fr_mul(&carry, &carry, &tail);
    }
}
/* src/shuffle.rs:142-209 */
void zo_shuffle_z(zo_fr *z, const zo_fr *const *inputs, const zo_fr *const *shuffles,
                  int group_len, const zo_fr *beta, size_t n) {
    zo_fr *bp = (zo_fr *)malloc(sizeof(zo_fr) * group_len);
    zo_fr *den = (zo_fr *)malloc(sizeof(zo_fr) * n);
    bp[0] = *beta;
    for (int j = 1; j < group_len; j++) fr_mul(&bp[j], &bp[j - 1], beta);
    for (size_t i = 0; i < n; i++) {
        zo_fr acc; fr_one(&acc);
        for (int j = 0; j < group_len; j++) {
            zo_fr t; fr_add(&t, &shuffles[j][i], &bp[j]); fr_mul(&acc, &acc, &t);
        }
        den[i] = acc;
    }
    zo_batch_invert(den, n);
    for (size_t i = 0; i < n; i++)
        for (int j = 0; j < group_len; j++) {
            zo_fr t; fr_add(&t, &inputs[j][i], &bp[j]); fr_mul(&den[i], &den[i], &t);
        }
    zo_fr acc; fr_one(&acc);
    for (size_t i = 0; i < n; i++) { z[i] = acc; fr_mul(&acc, &acc, &den[i]); }
    free(den); free(bp);
}

/* ---- eval_h per-row formulas --------------------------------------------- */
static inline size_t rot_idx(size_t i, long rot, size_t n) {
    long long v = ((long long)i + rot) % (long long)n;
    if (v < 0) v += (long long)n;
    return (size_t)v;
}
/* cuda/bn254.cu:585-639 (op 0 add, 1 mul, 2 copy, 3 sub); a NULL l takes l_c,
 * a NULL r takes r_c */
void zo_field_op(zo_fr *res, const zo_fr *l, int l_rot, const zo_fr *l_c,
                 const zo_fr *r, int r_rot, const zo_fr *r_c, size_t n, int op) {
    zo_fr *out = (zo_fr *)malloc(sizeof(zo_fr) * n);
    for (size_t i = 0; i < n; i++) {
        zo_fr fl = {{0}}, fr = {{0}};
        if (l) { fl = l[rot_idx(i, l_rot, n)]; if (l_c) fr_mul(&fl, &fl, l_c); }
        else fl = *l_c;
        if (r) { fr = r[rot_idx(i, r_rot, n)]; if (r_c) fr_mul(&fr, &fr, r_c); }
        else if (r_c) fr = *r_c;
        switch (op) {
        case 0: fr_add(&out[i], &fl, &fr); break;
        case 1: fr_mul(&out[i], &fl, &fr); break;
        case 2: out[i] = fl; break;
        default: fr_sub(&out[i], &fl, &fr); break;
        }
    }
    memcpy(res, out, sizeof(zo_fr) * n);
    free(out);
}
/* cuda/bn254.cu:542-572 */
void zo_field_op_batch_mul_sum(zo_fr *res, const zo_fr *const *vals, const int *rots,
                               const zo_fr *coeffs, const int *term_off, int nterms,
                               size_t n) {
#pragma omp parallel for OMP_THREADS if (n >= 4096)
    for (size_t i = 0; i < n; i++) {
        zo_fr acc = {{0}};
        for (int t = 0; t < nterms; t++) {
            zo_fr p = coeffs[t];
            for (int k = term_off[t]; k < term_off[t + 1]; k++)
                fr_mul(&p, &p, &vals[k][(i + (size_t)(long)rots[k] + n) & (n - 1)]);
            fr_add(&acc, &acc, &p);
        }
        fr_add(&res[i], &res[i], &acc);
    }
}
void zo_field_mul_zip(zo_fr *buf, const zo_fr *coeff, size_t coeff_n, size_t n) {
    for (size_t i = 0; i < n; i++) fr_mul(&buf[i], &buf[i], &coeff[i % coeff_n]);
}
/* cuda/bn254.cu:664-702 */
void zo_permutation_eval_h_p1(zo_fr *res, const zo_fr *first, const zo_fr *last,
                              const zo_fr *l0, const zo_fr *l_last, const zo_fr *y, size_t n) {
    zo_fr one; fr_one(&one);
    for (size_t i = 0; i < n; i++) {
        zo_fr t1 = res[i], t2;
        fr_mul(&t1, &t1, y);
        fr_sub(&t2, &one, &first[i]); fr_mul(&t2, &t2, &l0[i]); fr_add(&t1, &t1, &t2);
        fr_mul(&t1, &t1, y);
        fr_mul(&t2, &last[i], &last[i]); fr_sub(&t2, &t2, &last[i]); fr_mul(&t2, &t2, &l_last[i]);
        fr_add(&t1, &t1, &t2);
        res[i] = t1;
    }
}
/* cuda/bn254.cu:704-739 (also logup_eval_h_z_set, :1526-1540) */
void zo_permutation_eval_h_p2(zo_fr *res, const zo_fr *const *set, const zo_fr *l0,
                              const zo_fr *y, int n_set, int rot, size_t n) {
    for (size_t i = 0; i < n; i++) {
        size_t rp = (i + n + (size_t)(long)rot) & (n - 1);
        zo_fr t1 = res[i], t2;
        for (int j = 1; j < n_set; j++) {
            fr_mul(&t1, &t1, y);
            fr_sub(&t2, &set[j][i], &set[j - 1][rp]);
            fr_mul(&t2, &t2, &l0[i]);
            fr_add(&t1, &t1, &t2);
        }
        res[i] = t1;
    }
}
/* cuda/bn254.cu:741-765 */
void zo_permutation_eval_h_l(zo_fr *res, const zo_fr *beta, const zo_fr *gamma,
                             const zo_fr *p, size_t n) {
    for (size_t i = 0; i < n; i++) {
        zo_fr t; fr_mul(&t, &p[i], beta);
        if (i == 0) fr_add(&t, &t, gamma);
        fr_add(&res[i], &res[i], &t);
    }
}
/* cuda/bn254.cu:867-918 */
void zo_logup_eval_h(zo_fr *res, const zo_fr *input_product, const zo_fr *input_product_sum,
                     const zo_fr *table, const zo_fr *multiplicity, const zo_fr *z_first,
                     const zo_fr *z_last, const zo_fr *l0, const zo_fr *l_last,
                     const zo_fr *l_active_row, const zo_fr *y, int rot, size_t n) {
    for (size_t i = 0; i < n; i++) {
        size_t rn = (i + (size_t)(long)rot) & (n - 1);
        zo_fr t = res[i], u, p;
        fr_mul(&t, &t, y); fr_mul(&u, &l0[i], &z_first[i]); fr_add(&t, &t, &u);
        fr_mul(&t, &t, y); fr_mul(&u, &l_last[i], &z_last[i]); fr_add(&t, &t, &u);
        fr_mul(&t, &t, y);
        fr_sub(&u, &z_first[rn], &z_first[i]);
        fr_mul(&u, &u, &table[i]);
        fr_add(&u, &u, &multiplicity[i]);
        fr_mul(&u, &u, &input_product[i]);
        fr_mul(&p, &table[i], &input_product_sum[i]);
        fr_sub(&u, &u, &p);
        fr_mul(&u, &u, &l_active_row[i]);
        fr_add(&t, &t, &u);
        res[i] = t;
    }
}
/* cuda/bn254.cu:920-949 */
void zo_logup_eval_h_extra_inputs(zo_fr *res, const zo_fr *input_product,
                                  const zo_fr *input_product_sum, const zo_fr *z,
                                  const zo_fr *l_active_row, const zo_fr *y, int rot, size_t n) {
    for (size_t i = 0; i < n; i++) {
        size_t rn = (i + (size_t)(long)rot) & (n - 1);
        zo_fr t = res[i], u;
        fr_mul(&t, &t, y);
        fr_sub(&u, &z[rn], &z[i]);
        fr_mul(&u, &u, &input_product[i]);
        fr_sub(&u, &u, &input_product_sum[i]);
        fr_mul(&u, &u, &l_active_row[i]);
        fr_add(&t, &t, &u);
        res[i] = t;
    }
}
/* cuda/bn254.cu:496-540: res[i] = sum_j v_j[i] w^(rot_j i mod n) (* c_j) */
void zo_field_sum(zo_fr *res, const zo_fr *const *v, const zo_fr *const *v_c, const int *v_rot,
                  const zo_fr *omegas, int v_n, size_t n) {
    for (size_t i = 0; i < n; i++) {
        zo_fr acc = {{0}};
        for (int j = 0; j < v_n; j++) {
            long long r = ((long long)v_rot[j] % (long long)n + (long long)n) % (long long)n;
            size_t e = (size_t)(((unsigned long long)r * i) % n);
            zo_fr t;
            fr_mul(&t, &v[j][i], &omegas[e]);
            if (v_c[j]) fr_mul(&t, &t, v_c[j]);
            if (j == 0) acc = t; else fr_add(&acc, &acc, &t);
        }
        res[i] = acc;
    }
}
/* cuda/bn254.cu:798-866 */
void zo_lookup_eval_h(zo_fr *res, const zo_fr *input, const zo_fr *table, const zo_fr *permuted_input,
                      const zo_fr *permuted_table, const zo_fr *z, const zo_fr *l0, const zo_fr *l_last,
                      const zo_fr *l_active_row, const zo_fr *y, const zo_fr *beta, const zo_fr *gamma,
                      int rot, size_t n) {
    zo_fr one; fr_one(&one);
    for (size_t i = 0; i < n; i++) {
        size_t rn = (i + (size_t)(long)rot) & (n - 1);
        size_t rp = (i + n - (size_t)(long)rot) & (n - 1);
        zo_fr t = res[i], u, p, x;
        fr_mul(&t, &t, y); fr_sub(&u, &one, &z[i]); fr_mul(&u, &l0[i], &u); fr_add(&t, &t, &u);
        fr_mul(&t, &t, y); fr_mul(&u, &z[i], &z[i]); fr_sub(&u, &u, &z[i]); fr_mul(&u, &l_last[i], &u);
        fr_add(&t, &t, &u);
        fr_mul(&t, &t, y);
        fr_add(&u, &permuted_input[i], beta);
        fr_add(&p, &permuted_table[i], gamma);
        fr_mul(&u, &u, &p);
        fr_mul(&u, &u, &z[rn]);
        fr_mul(&x, &input[i], &table[i]);
        fr_mul(&x, &z[i], &x);
        fr_sub(&u, &u, &x);
        fr_mul(&u, &u, &l_active_row[i]);
        fr_add(&t, &t, &u);
        fr_mul(&t, &t, y);
        fr_sub(&p, &permuted_input[i], &permuted_table[i]);
        fr_mul(&u, &l0[i], &p);
        fr_add(&t, &t, &u);
        fr_mul(&t, &t, y);
        fr_sub(&u, &permuted_input[i], &permuted_input[rp]);
        fr_mul(&u, &u, &p);
        fr_mul(&u, &u, &l_active_row[i]);
        fr_add(&t, &t, &u);
        res[i] = t;
    }
}
/* cuda/bn254.cu:1172-1213 */
void zo_logup_eval_h_inputs_product_sum(zo_fr *product, zo_fr *product_sum,
                                        const zo_fr *const *set, int n_set, size_t n) {
    for (size_t i = 0; i < n; i++) {
        zo_fr p; fr_one(&p);
        for (int j = 0; j < n_set; j++) fr_mul(&p, &p, &set[j][i]);
        product[i] = p;
        zo_fr sum = {{0}};
        for (int j = 0; j < n_set; j++) {
            zo_fr u; fr_one(&u);
            for (int k = 0; k < n_set; k++)
                if (k != j) fr_mul(&u, &u, &set[k][i]);
            fr_add(&sum, &sum, &u);
        }
        product_sum[i] = sum;
    }
}
/* cuda/bn254.cu:975-1016 */
void zo_shuffle_eval_h(zo_fr *res, const zo_fr *input, const zo_fr *table, const zo_fr *z,
                       const zo_fr *l0, const zo_fr *l_last, const zo_fr *l_active_row,
                       const zo_fr *y, int rot, size_t n) {
    zo_fr one; fr_one(&one);
    for (size_t i = 0; i < n; i++) {
        size_t rn = (i + (size_t)(long)rot) & (n - 1);
        zo_fr t = res[i], u, w;
        fr_mul(&t, &t, y); fr_sub(&u, &one, &z[i]); fr_mul(&u, &l0[i], &u); fr_add(&t, &t, &u);
        fr_mul(&t, &t, y); fr_mul(&u, &z[i], &z[i]); fr_sub(&u, &u, &z[i]); fr_mul(&u, &l_last[i], &u);
        fr_add(&t, &t, &u);
        fr_mul(&t, &t, y);
        fr_mul(&u, &table[i], &z[rn]);
        fr_mul(&w, &input[i], &z[i]);
        fr_sub(&u, &u, &w);
        fr_mul(&u, &u, &l_active_row[i]);
        fr_add(&t, &t, &u);
        res[i] = t;
    }
}
/* cuda/bn254.cu:1018-1071 */
void zo_shuffle_eval_h_v2(zo_fr *res, const zo_fr *const *inputs, const zo_fr *const *tables,
                          const zo_fr *betas, int group_len, const zo_fr *z, const zo_fr *l0,
                          const zo_fr *l_last, const zo_fr *l_active_row, const zo_fr *y,
                          int rot, size_t n) {
    zo_fr one; fr_one(&one);
    for (size_t i = 0; i < n; i++) {
        size_t rn = (i + (size_t)(long)rot) & (n - 1);
        zo_fr l, r, u, w;
        fr_add(&l, &inputs[0][i], &betas[0]);
        fr_add(&r, &tables[0][i], &betas[0]);
        for (int j = 1; j < group_len; j++) {
            fr_add(&u, &inputs[j][i], &betas[j]); fr_mul(&l, &l, &u);
            fr_add(&u, &tables[j][i], &betas[j]); fr_mul(&r, &r, &u);
        }
        zo_fr t = res[i];
        fr_mul(&t, &t, y); fr_sub(&u, &one, &z[i]); fr_mul(&u, &l0[i], &u); fr_add(&t, &t, &u);
        fr_mul(&t, &t, y); fr_mul(&u, &z[i], &z[i]); fr_sub(&u, &u, &z[i]); fr_mul(&u, &l_last[i], &u);
        fr_add(&t, &t, &u);
        fr_mul(&t, &t, y);
        fr_mul(&u, &r, &z[rn]);
        fr_mul(&w, &l, &z[i]);
        fr_sub(&u, &u, &w);
        fr_mul(&u, &u, &l_active_row[i]);
        fr_add(&t, &t, &u);
        res[i] = t;
    }
}

/* ---- multi-open ------------------------------------------------------------ */
/* cuda/bn254.cu:1608-1633 with _poly_eval :238-256: rounds of at most 4 bits;
 * a round with step 2^deg folds each group of `count` = 2^cur entries at
 * stride 2^deg by Horner in x_table[deg] = x^(2^deg) into its first entry.
 * res (n entries) is the work area; res[0] = p(x) at the end (n >= 2; for
 * n = 1 the reference writes nothing, here res[0] = p[0]). */
void zo_poly_eval(zo_fr *res, const zo_fr *p, const zo_fr *x_table, size_t n) {
    if (n == 1) { res[0] = p[0]; return; }
    int deg = 0, remain = __builtin_ctzll(n);
    size_t m = n;
    while (m > 1) {
        int cur = remain < 4 ? remain : 4;
        size_t count = (size_t)1 << cur, groups = m >> cur, step = (size_t)1 << deg;
        const zo_fr *src = deg == 0 ? p : res;
        zo_fr x = x_table[deg];
        for (size_t g = 0; g < groups; g++) {
            size_t start = (g * count) << deg;
            size_t end = start + ((count - 1) << deg);
            zo_fr acc = src[end];
            for (size_t i = end - step; i > start; i -= step) {
                fr_mul(&acc, &acc, &x);
                fr_add(&acc, &acc, &src[i]);
            }
            fr_mul(&acc, &acc, &x);
            fr_add(&res[start], &acc, &src[start]);
        }
        m >>= cur;
        deg += cur;
        remain -= cur;
    }
}
/* src/multiopen.rs:145-159 (GWC) and :543-553 (Shplonk): in-place division
 * by (X - z): poly[i] = q_i for i <= n-2, poly[n-1] = 0 (the remainder, and so
 * poly[0] before the call, is dropped). */
void zo_kate_divide(zo_fr *poly, const zo_fr *z, size_t n) {
    if (n < 2) return; /* the loop leaves a single entry as it is */
    zo_fr tmp = poly[n - 1];
    memset(&poly[n - 1], 0, sizeof(zo_fr));
    for (size_t i = n - 2; i >= 1; i--) {
        zo_fr t, q;
        fr_mul(&t, &tmp, z);
        fr_add(&q, &poly[i], &t);
        poly[i] = tmp;
        tmp = q;
    }
    poly[0] = tmp;
}
/* cuda/bn254.cu:1133-1151 */
void zo_shplonk_h_x_merge(zo_fr *res, const zo_fr *v, const zo_fr *values, const zo_fr *omegas,
                          const zo_fr *diff, int diff_n, size_t n) {
#pragma omp parallel for OMP_THREADS if (n >= 4096)
    for (size_t i = 0; i < n; i++) {
        zo_fr t = values[i];
        for (int j = 0; j < diff_n; j++) {
            zo_fr d; fr_sub(&d, &omegas[i], &diff[j]);
            fr_mul(&t, &t, &d);
        }
        zo_fr r; fr_mul(&r, &res[i], v);
        fr_add(&res[i], &r, &t);
    }
}
/* cuda/bn254.cu:1153-1170: values[i] *= 1 / prod_j (omegas[i] - points[j]);
 * the reference asserts the product is nonzero, here a zero product gives 0 */
void zo_shplonk_h_x_div_points(zo_fr *values, const zo_fr *omegas, const zo_fr *points,
                               int points_n, size_t n) {
#pragma omp parallel for OMP_THREADS if (n >= 1024)
    for (size_t i = 0; i < n; i++) {
        zo_fr t; fr_sub(&t, &omegas[i], &points[0]);
        for (int j = 1; j < points_n; j++) {
            zo_fr d; fr_sub(&d, &omegas[i], &points[j]);
            fr_mul(&t, &t, &d);
        }
        zo_fr inv; zo_fr_inv(&inv, &t);
        fr_mul(&values[i], &values[i], &inv);
    }
}
