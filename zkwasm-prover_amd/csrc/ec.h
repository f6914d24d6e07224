// BN254 G1 (y^2 = x^3 + 3 over Fp) for gfx950.
//
// Storage formats match the reference C-ABI (cuda/zprize_ec_wrapper.cuh:72-73):
//   G1Aff  {x, y}     64 B, identity = (0, 0)   (halo2 G1Affine)
//   G1Jac  {x, y, z}  96 B, identity has z = 0  (halo2 G1, Jacobian)
// Internally the MSM accumulates in XYZZ coordinates (x = X/ZZ, y = Y/ZZZ,
// ZZ^3 = ZZZ^2): mixed add 8M+2S, add 12M+2S, dbl 6M+3S, with no
// inversion and simple exceptional cases.  The Horner tail uses Jacobian
// doubling (dbl-2009-l, 2M+5S).
#pragma once
#include "ff.h"

namespace zk {

struct alignas(32) G1Aff { Fp x, y; };
struct alignas(32) G1Jac { Fp x, y, z; };
struct alignas(32) G1Xyzz { Fp x, y, zz, zzz; };

__device__ __forceinline__ bool aff_is_id(const G1Aff &p) {
    return fe_is_zero<FpParams>(p.x) && fe_is_zero<FpParams>(p.y);
}
__device__ __forceinline__ G1Xyzz xyzz_id() {
    G1Xyzz r;
    r.x = fe_one<FpParams>();
    r.y = fe_one<FpParams>();
    r.zz = fe_zero<FpParams>();
    r.zzz = fe_zero<FpParams>();
    return r;
}
__device__ __forceinline__ bool xyzz_is_id(const G1Xyzz &p) { return fe_is_zero<FpParams>(p.zz); }
__device__ __forceinline__ G1Xyzz xyzz_from_aff(const G1Aff &p, bool neg) {
    if (aff_is_id(p)) return xyzz_id();
    G1Xyzz r;
    r.x = p.x;
    r.y = neg ? fe_neg<FpParams>(p.y) : p.y;
    r.zz = fe_one<FpParams>();
    r.zzz = fe_one<FpParams>();
    return r;
}

// dbl-2008-s-1 (a = 0)
__device__ __forceinline__ G1Xyzz xyzz_dbl(const G1Xyzz &p) {
    if (xyzz_is_id(p)) return p;
    Fp U = fe_dbl<FpParams>(p.y);
    Fp V = fe_sqr<FpParams>(U);
    Fp W = U * V;
    Fp S = p.x * V;
    Fp X2 = fe_sqr<FpParams>(p.x);
    Fp M = fe_add<FpParams>(fe_dbl<FpParams>(X2), X2);
    G1Xyzz r;
    r.x = fe_sub<FpParams>(fe_sqr<FpParams>(M), fe_dbl<FpParams>(S));
    r.y = fe_sub<FpParams>(M * fe_sub<FpParams>(S, r.x), W * p.y);
    r.zz = V * p.zz;
    r.zzz = W * p.zzz;
    return r;
}

// madd-2008-s: p + (x2, y2) with y2 already negated by the caller if needed
__device__ __forceinline__ G1Xyzz xyzz_add_aff(const G1Xyzz &p, const Fp &x2, const Fp &y2) {
    if (xyzz_is_id(p)) {
        G1Xyzz r;
        r.x = x2;
        r.y = y2;
        r.zz = fe_one<FpParams>();
        r.zzz = fe_one<FpParams>();
        return r;
    }
    Fp U2 = x2 * p.zz;
    Fp S2 = y2 * p.zzz;
    Fp P = fe_sub<FpParams>(U2, p.x);
    Fp R = fe_sub<FpParams>(S2, p.y);
    if (fe_is_zero<FpParams>(P)) {
        if (fe_is_zero<FpParams>(R)) {
            G1Xyzz a;
            a.x = x2; a.y = y2; a.zz = fe_one<FpParams>(); a.zzz = fe_one<FpParams>();
            return xyzz_dbl(a);
        }
        return xyzz_id();
    }
    Fp PP = fe_sqr<FpParams>(P);
    Fp PPP = P * PP;
    Fp Q = p.x * PP;
    G1Xyzz r;
    r.x = fe_sub<FpParams>(fe_sub<FpParams>(fe_sqr<FpParams>(R), PPP), fe_dbl<FpParams>(Q));
    r.y = fe_sub<FpParams>(R * fe_sub<FpParams>(Q, r.x), p.y * PPP);
    r.zz = p.zz * PP;
    r.zzz = p.zzz * PPP;
    return r;
}

// add-2008-s
__device__ __forceinline__ G1Xyzz xyzz_add(const G1Xyzz &p, const G1Xyzz &q) {
    if (xyzz_is_id(p)) return q;
    if (xyzz_is_id(q)) return p;
    Fp U1 = p.x * q.zz;
    Fp U2 = q.x * p.zz;
    Fp S1 = p.y * q.zzz;
    Fp S2 = q.y * p.zzz;
    Fp P = fe_sub<FpParams>(U2, U1);
    Fp R = fe_sub<FpParams>(S2, S1);
    if (fe_is_zero<FpParams>(P)) {
        if (fe_is_zero<FpParams>(R)) return xyzz_dbl(p);
        return xyzz_id();
    }
    Fp PP = fe_sqr<FpParams>(P);
    Fp PPP = P * PP;
    Fp Q = U1 * PP;
    G1Xyzz r;
    r.x = fe_sub<FpParams>(fe_sub<FpParams>(fe_sqr<FpParams>(R), PPP), fe_dbl<FpParams>(Q));
    r.y = fe_sub<FpParams>(R * fe_sub<FpParams>(Q, r.x), S1 * PPP);
    r.zz = p.zz * q.zz * PP;
    r.zzz = p.zzz * q.zzz * PPP;
    return r;
}

// XYZZ -> Jacobian without inversion: Z = ZZ*ZZZ, X = X*ZZ*ZZZ^2, Y = Y*ZZZ^4
// (x = X/Z^2, y = Y/Z^3 using ZZ^3 = ZZZ^2).
__device__ __forceinline__ G1Jac xyzz_to_jac(const G1Xyzz &p) {
    G1Jac r;
    if (xyzz_is_id(p)) {
        r.x = fe_zero<FpParams>();
        r.y = fe_one<FpParams>();
        r.z = fe_zero<FpParams>();
        return r;
    }
    Fp zzz2 = fe_sqr<FpParams>(p.zzz);
    r.z = p.zz * p.zzz;
    r.x = p.x * p.zz * zzz2;
    r.y = p.y * fe_sqr<FpParams>(zzz2);
    return r;
}

// Jacobian dbl-2009-l (a = 0)
__device__ __forceinline__ G1Jac jac_dbl(const G1Jac &p) {
    if (fe_is_zero<FpParams>(p.z)) return p;
    Fp A = fe_sqr<FpParams>(p.x);
    Fp B = fe_sqr<FpParams>(p.y);
    Fp C = fe_sqr<FpParams>(B);
    Fp D = fe_dbl<FpParams>(fe_sub<FpParams>(fe_sub<FpParams>(fe_sqr<FpParams>(fe_add<FpParams>(p.x, B)), A), C));
    Fp E = fe_add<FpParams>(fe_dbl<FpParams>(A), A);
    Fp F = fe_sqr<FpParams>(E);
    G1Jac r;
    r.x = fe_sub<FpParams>(F, fe_dbl<FpParams>(D));
    Fp C8 = fe_dbl<FpParams>(fe_dbl<FpParams>(fe_dbl<FpParams>(C)));
    r.y = fe_sub<FpParams>(E * fe_sub<FpParams>(D, r.x), C8);
    r.z = fe_dbl<FpParams>(p.y * p.z);
    return r;
}

// Jacobian add-2007-bl with the equal / opposite / identity cases
__device__ __forceinline__ G1Jac jac_add(const G1Jac &p, const G1Jac &q) {
    if (fe_is_zero<FpParams>(p.z)) return q;
    if (fe_is_zero<FpParams>(q.z)) return p;
    Fp Z1Z1 = fe_sqr<FpParams>(p.z);
    Fp Z2Z2 = fe_sqr<FpParams>(q.z);
    Fp U1 = p.x * Z2Z2;
    Fp U2 = q.x * Z1Z1;
    Fp S1 = p.y * q.z * Z2Z2;
    Fp S2 = q.y * p.z * Z1Z1;
    if (fe_eq<FpParams>(U1, U2)) {
        if (fe_eq<FpParams>(S1, S2)) return jac_dbl(p);
        G1Jac r;
        r.x = fe_zero<FpParams>();
        r.y = fe_one<FpParams>();
        r.z = fe_zero<FpParams>();
        return r;
    }
    Fp H = fe_sub<FpParams>(U2, U1);
    Fp I = fe_sqr<FpParams>(fe_dbl<FpParams>(H));
    Fp J = H * I;
    Fp rr = fe_dbl<FpParams>(fe_sub<FpParams>(S2, S1));
    Fp V = U1 * I;
    G1Jac r;
    r.x = fe_sub<FpParams>(fe_sub<FpParams>(fe_sqr<FpParams>(rr), J), fe_dbl<FpParams>(V));
    r.y = fe_sub<FpParams>(rr * fe_sub<FpParams>(V, r.x), fe_dbl<FpParams>(S1 * J));
    Fp zs = fe_add<FpParams>(p.z, q.z);
    r.z = fe_sub<FpParams>(fe_sub<FpParams>(fe_sqr<FpParams>(zs), Z1Z1), Z2Z2) * H;
    return r;
}

// Normalise to z = 1.  The identity is emitted as (0, 1*R, 0), the
// reference's Jacobian identity (cuda/zprize/ec.cuh:142).
__device__ __forceinline__ G1Jac jac_normalize(const G1Jac &p) {
    G1Jac r;
    if (fe_is_zero<FpParams>(p.z)) {
        r.x = fe_zero<FpParams>();
        r.y = fe_one<FpParams>();
        r.z = fe_zero<FpParams>();
        return r;
    }
    Fp zi = fe_inv<FpParams>(p.z);
    Fp zi2 = fe_sqr<FpParams>(zi);
    r.x = p.x * zi2;
    r.y = p.y * zi2 * zi;
    r.z = fe_one<FpParams>();
    return r;
}

__device__ __forceinline__ G1Aff aff_load(const G1Aff *p) {
    G1Aff r;
    r.x = fe_load<FpParams>(&p->x);
    r.y = fe_load<FpParams>(&p->y);
    return r;
}
__device__ __forceinline__ G1Xyzz xyzz_load(const G1Xyzz *p) {
    G1Xyzz r;
    r.x = fe_load<FpParams>(&p->x);
    r.y = fe_load<FpParams>(&p->y);
    r.zz = fe_load<FpParams>(&p->zz);
    r.zzz = fe_load<FpParams>(&p->zzz);
    return r;
}
__device__ __forceinline__ void xyzz_store(G1Xyzz *p, const G1Xyzz &r) {
    fe_store<FpParams>(&p->x, r.x);
    fe_store<FpParams>(&p->y, r.y);
    fe_store<FpParams>(&p->zz, r.zz);
    fe_store<FpParams>(&p->zzz, r.zzz);
}
__device__ __forceinline__ void jac_store(G1Jac *p, const G1Jac &r) {
    fe_store<FpParams>(&p->x, r.x);
    fe_store<FpParams>(&p->y, r.y);
    fe_store<FpParams>(&p->z, r.z);
}
__device__ __forceinline__ G1Jac jac_load(const G1Jac *p) {
    G1Jac r;
    r.x = fe_load<FpParams>(&p->x);
    r.y = fe_load<FpParams>(&p->y);
    r.z = fe_load<FpParams>(&p->z);
    return r;
}

}  // namespace zk
