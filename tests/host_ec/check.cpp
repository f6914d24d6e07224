// CPU harness for the host curve arithmetic of the MSM tail (csrc/host_ec.h).
#include <string.h>

#include "host_ec.h"

using namespace zkh;

extern "C" {
void he_add(void *r, const void *a, const void *b) {
    Fp x = fp_add(*(const Fp *)a, *(const Fp *)b);
    memcpy(r, &x, 32);
}
void he_sub(void *r, const void *a, const void *b) {
    Fp x = fp_sub(*(const Fp *)a, *(const Fp *)b);
    memcpy(r, &x, 32);
}
void he_mul(void *r, const void *a, const void *b) {
    Fp x = fp_mul(*(const Fp *)a, *(const Fp *)b);
    memcpy(r, &x, 32);
}
void he_inv(void *r, const void *a) {
    Fp x = fp_inv(*(const Fp *)a);
    memcpy(r, &x, 32);
}
// Jacobian (x, y, z) in Montgomery form, 96 bytes
void he_jac_add(void *r, const void *p, const void *q) {
    Jac x = jac_add(*(const Jac *)p, *(const Jac *)q);
    memcpy(r, &x, 96);
}
void he_jac_dbl(void *r, const void *p) {
    Jac x = jac_dbl(*(const Jac *)p);
    memcpy(r, &x, 96);
}
void he_jac_normalize(void *r, const void *p) {
    Jac x = jac_normalize(*(const Jac *)p);
    memcpy(r, &x, 96);
}
void he_from_xyzz(void *r, const void *xyzz) {
    const Fp *f = (const Fp *)xyzz;
    Jac x = jac_from_xyzz(f[0], f[1], f[2], f[3]);
    memcpy(r, &x, 96);
}
}
extern "C" {
// the MSM tail over W * (nb + 1) XYZZ masked sums (128 B each), window-major
void he_msm_tail(void *r, const void *q, uint32_t c, uint32_t W, uint32_t nb, uint32_t l1) {
    Jac x = msm_tail((const Xyzz *)q, c, W, nb, l1);
    memcpy(r, &x, 96);
}
// the same sum split at exponent lo: 2^lo (exponents lo..top) + (exponents 0..lo-1)
void he_msm_tail_split(void *r, const void *q, uint32_t c, uint32_t W, uint32_t nb, uint32_t l1, int lo) {
    const Xyzz *p = (const Xyzz *)q;
    Xyzz a = msm_tail_range(p, c, W, nb, l1, lo, msm_tail_top(c, W, nb, l1));
    for (int k = 0; k < lo; k++) a = xyzz_dbl(a);
    Jac x = xyzz_normalize(xyzz_add(a, msm_tail_range(p, c, W, nb, l1, 0, lo - 1)));
    memcpy(r, &x, 96);
}
}
