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
