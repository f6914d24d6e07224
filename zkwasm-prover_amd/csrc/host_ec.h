// Host-side BN254 Fp / G1 Jacobian arithmetic for the serial MSM tail
// (window Horner).  Same byte layout as the device types: four
// little-endian u64 limbs in Montgomery form (R = 2^256).
#pragma once
#include <stdint.h>
#include <string.h>

#include "inv_bingcd.h"

namespace zkh {

typedef unsigned __int128 u128;

struct Fp {
    uint64_t l[4];
};
struct Jac {
    Fp x, y, z;
};

constexpr uint64_t kP[4] = {0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL,
                            0x30644e72e131a029ULL};
constexpr uint64_t kPInv = 0x87d20782e4866389ULL;  // -p^-1 mod 2^64
constexpr uint64_t kOne[4] = {0xd35d438dc58f0d9dULL, 0x0a78eb28f5c70b3dULL, 0x666ea36f7879462cULL,
                              0x0e0a77c19a07df2fULL};  // R mod p

// word-wise add / subtract with carry (compile to add-with-carry chains)
inline uint64_t addc(uint64_t a, uint64_t b, uint64_t &c) {
    const u128 t = (u128)a + b + c;
    c = (uint64_t)(t >> 64);
    return (uint64_t)t;
}
inline uint64_t subb(uint64_t a, uint64_t b, uint64_t &br) {
    const u128 t = (u128)a - b - br;
    br = (uint64_t)(t >> 64) & 1;
    return (uint64_t)t;
}
// a - p if a >= p (a < 2^257 as 4 words plus the carry-out `hi`), else a:
// both values formed and one selected by mask (the tail's values are random,
// so data-dependent branches mispredict half the time)
inline void reduce_once(uint64_t a[4], uint64_t hi) {
    uint64_t d[4], br = 0;
    for (int i = 0; i < 4; i++) d[i] = subb(a[i], kP[i], br);
    const uint64_t keep = (uint64_t)0 - (br & (hi ^ 1));  // a < p
    for (int i = 0; i < 4; i++) a[i] = (a[i] & keep) | (d[i] & ~keep);
}
inline bool geq_p(const uint64_t a[4]) {
    uint64_t br = 0;
    for (int i = 0; i < 4; i++) (void)subb(a[i], kP[i], br);
    return br == 0;
}
inline void sub_p(uint64_t a[4]) {
    uint64_t br = 0;
    for (int i = 0; i < 4; i++) a[i] = subb(a[i], kP[i], br);
}
inline Fp fp_zero() { return Fp{{0, 0, 0, 0}}; }
inline Fp fp_one() { return Fp{{kOne[0], kOne[1], kOne[2], kOne[3]}}; }
inline bool fp_is_zero(const Fp &a) { return !(a.l[0] | a.l[1] | a.l[2] | a.l[3]); }
inline bool fp_eq(const Fp &a, const Fp &b) { return memcmp(&a, &b, sizeof(Fp)) == 0; }

inline Fp fp_add(const Fp &a, const Fp &b) {
    Fp r;
    uint64_t c = 0;
    for (int i = 0; i < 4; i++) r.l[i] = addc(a.l[i], b.l[i], c);
    reduce_once(r.l, c);
    return r;
}
inline Fp fp_sub(const Fp &a, const Fp &b) {
    Fp r;
    uint64_t br = 0;
    for (int i = 0; i < 4; i++) r.l[i] = subb(a.l[i], b.l[i], br);
    // p added back under a mask when the difference went negative
    const uint64_t m = (uint64_t)0 - br, c0 = 0;
    uint64_t c = c0;
    for (int i = 0; i < 4; i++) r.l[i] = addc(r.l[i], kP[i] & m, c);
    return r;
}
// CIOS Montgomery product on 64-bit limbs
inline Fp fp_mul(const Fp &a, const Fp &b) {
    uint64_t t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; i++) {
        u128 c = 0;
        for (int j = 0; j < 4; j++) {
            c = (u128)a.l[j] * b.l[i] + t[j] + (uint64_t)(c >> 64);
            t[j] = (uint64_t)c;
        }
        c = (u128)t[4] + (uint64_t)(c >> 64);
        t[4] = (uint64_t)c;
        t[5] = (uint64_t)(c >> 64);
        uint64_t m = t[0] * kPInv;
        c = (u128)m * kP[0] + t[0];
        for (int j = 1; j < 4; j++) {
            c = (u128)m * kP[j] + t[j] + (uint64_t)(c >> 64);
            t[j - 1] = (uint64_t)c;
        }
        c = (u128)t[4] + (uint64_t)(c >> 64);
        t[3] = (uint64_t)c;
        t[4] = t[5] + (uint64_t)(c >> 64);
    }
    Fp r{{t[0], t[1], t[2], t[3]}};
    reduce_once(r.l, t[4]);
    return r;
}
inline Fp fp_sqr(const Fp &a) { return fp_mul(a, a); }
inline Fp fp_dbl(const Fp &a) { return fp_add(a, a); }
// a^(p-2) (Fermat); 0 -> 0
inline Fp fp_inv_fermat(const Fp &a) {
    uint64_t e[4] = {kP[0] - 2, kP[1], kP[2], kP[3]};
    Fp acc = fp_one(), base = a;
    for (int i = 0; i < 4; i++)
        for (int b = 0; b < 64; b++) {
            if ((e[i] >> b) & 1) acc = fp_mul(acc, base);
            base = fp_sqr(base);
        }
    return acc;
}
// Montgomery inverse by the optimized binary GCD (inv_bingcd.h): the cofactor
// v of y = aR times 2^36 R^3 under one Montgomery product is a^-1 R; checked
// by one product, Fermat otherwise.  0 -> 0.
inline Fp fp_inv(const Fp &a) {
    if (fp_is_zero(a)) return fp_zero();
    static const uint32_t kP32[8] = {0xd87cfd47u, 0x3c208c16u, 0x6871ca8du, 0x97816a91u,
                                     0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
    static const Fp kInvC = {{0x12f8df1fc57e1fafULL, 0x516b8b3ca75c06e1ULL, 0x448dbc80e6486ed4ULL,
                              0x00d762b62a2a6c68ULL}};  // 2^36 R^3 mod p
    uint32_t y[8], v[8];
    for (int i = 0; i < 4; i++) {
        y[2 * i] = (uint32_t)a.l[i];
        y[2 * i + 1] = (uint32_t)(a.l[i] >> 32);
    }
    const bool ok = zkinv::bingcd(v, y, kP32, 0xe4866389u);
    Fp r;
    for (int i = 0; i < 4; i++) r.l[i] = (uint64_t)v[2 * i] | ((uint64_t)v[2 * i + 1] << 32);
    r = fp_mul(r, kInvC);
    if (ok && fp_eq(fp_mul(a, r), fp_one())) return r;
    return fp_inv_fermat(a);
}

inline Jac jac_id() { return Jac{fp_zero(), fp_one(), fp_zero()}; }

// dbl-2009-l (a = 0)
inline Jac jac_dbl(const Jac &p) {
    if (fp_is_zero(p.z)) return p;
    Fp A = fp_sqr(p.x), B = fp_sqr(p.y), C = fp_sqr(B);
    Fp D = fp_dbl(fp_sub(fp_sub(fp_sqr(fp_add(p.x, B)), A), C));
    Fp E = fp_add(fp_dbl(A), A);
    Fp F = fp_sqr(E);
    Jac r;
    r.x = fp_sub(F, fp_dbl(D));
    r.y = fp_sub(fp_mul(E, fp_sub(D, r.x)), fp_dbl(fp_dbl(fp_dbl(C))));
    r.z = fp_dbl(fp_mul(p.y, p.z));
    return r;
}
// add-2007-bl with the identity / equal / opposite cases
inline Jac jac_add(const Jac &p, const Jac &q) {
    if (fp_is_zero(p.z)) return q;
    if (fp_is_zero(q.z)) return p;
    Fp Z1Z1 = fp_sqr(p.z), Z2Z2 = fp_sqr(q.z);
    Fp U1 = fp_mul(p.x, Z2Z2), U2 = fp_mul(q.x, Z1Z1);
    Fp S1 = fp_mul(fp_mul(p.y, q.z), Z2Z2), S2 = fp_mul(fp_mul(q.y, p.z), Z1Z1);
    if (fp_eq(U1, U2)) return fp_eq(S1, S2) ? jac_dbl(p) : jac_id();
    Fp H = fp_sub(U2, U1);
    Fp I = fp_sqr(fp_dbl(H));
    Fp J = fp_mul(H, I);
    Fp rr = fp_dbl(fp_sub(S2, S1));
    Fp V = fp_mul(U1, I);
    Jac r;
    r.x = fp_sub(fp_sub(fp_sqr(rr), J), fp_dbl(V));
    r.y = fp_sub(fp_mul(rr, fp_sub(V, r.x)), fp_dbl(fp_mul(S1, J)));
    r.z = fp_mul(fp_sub(fp_sub(fp_sqr(fp_add(p.z, q.z)), Z1Z1), Z2Z2), H);
    return r;
}
// XYZZ (x = X/ZZ, y = Y/ZZZ) -> Jacobian: Z = ZZ*ZZZ, X = X*ZZ*ZZZ^2, Y = Y*ZZZ^4
inline Jac jac_from_xyzz(const Fp &X, const Fp &Y, const Fp &ZZ, const Fp &ZZZ) {
    if (fp_is_zero(ZZ)) return jac_id();
    Fp zzz2 = fp_sqr(ZZZ);
    Jac r;
    r.z = fp_mul(ZZ, ZZZ);
    r.x = fp_mul(fp_mul(X, ZZ), zzz2);
    r.y = fp_mul(Y, fp_sqr(zzz2));
    return r;
}
// z = 1 form; the identity as (0, 1, 0)
inline Jac jac_normalize(const Jac &p) {
    if (fp_is_zero(p.z)) return jac_id();
    Fp zi = fp_inv(p.z), zi2 = fp_sqr(zi);
    return Jac{fp_mul(p.x, zi2), fp_mul(fp_mul(p.y, zi2), zi), fp_one()};
}

// XYZZ coordinates (x = X/ZZ, y = Y/ZZZ, ZZ^3 = ZZZ^2); ZZ = 0 is the identity.
// The device's masked sums arrive in this form, so the tail adds them with no
// conversion, and its additions cost 14 products against 16 + 5 for a
// Jacobian addition of a converted point.
struct Xyzz {
    Fp x, y, zz, zzz;
};
inline Xyzz xyzz_id() { return Xyzz{fp_one(), fp_one(), fp_zero(), fp_zero()}; }

// dbl-2008-s-1 (a = 0): 6M + 3S
inline Xyzz xyzz_dbl(const Xyzz &p) {
    if (fp_is_zero(p.zz)) return p;
    const Fp U = fp_dbl(p.y), V = fp_sqr(U), W = fp_mul(U, V), S = fp_mul(p.x, V);
    const Fp X2 = fp_sqr(p.x), M = fp_add(fp_dbl(X2), X2);
    Xyzz r;
    r.x = fp_sub(fp_sqr(M), fp_dbl(S));
    r.y = fp_sub(fp_mul(M, fp_sub(S, r.x)), fp_mul(W, p.y));
    r.zz = fp_mul(V, p.zz);
    r.zzz = fp_mul(W, p.zzz);
    return r;
}
// add-2008-s with the identity / equal / opposite cases: 12M + 2S
inline Xyzz xyzz_add(const Xyzz &p, const Xyzz &q) {
    if (fp_is_zero(p.zz)) return q;
    if (fp_is_zero(q.zz)) return p;
    const Fp U1 = fp_mul(p.x, q.zz), U2 = fp_mul(q.x, p.zz);
    const Fp S1 = fp_mul(p.y, q.zzz), S2 = fp_mul(q.y, p.zzz);
    const Fp P = fp_sub(U2, U1), R = fp_sub(S2, S1);
    if (fp_is_zero(P)) return fp_is_zero(R) ? xyzz_dbl(p) : xyzz_id();
    const Fp PP = fp_sqr(P), PPP = fp_mul(P, PP), Q = fp_mul(U1, PP);
    Xyzz r;
    r.x = fp_sub(fp_sub(fp_sqr(R), PPP), fp_dbl(Q));
    r.y = fp_sub(fp_mul(R, fp_sub(Q, r.x)), fp_mul(S1, PPP));
    r.zz = fp_mul(fp_mul(p.zz, q.zz), PP);
    r.zzz = fp_mul(fp_mul(p.zzz, q.zzz), PPP);
    return r;
}
// z = 1 Jacobian form (the identity as (0, 1, 0)): one inversion of ZZ * ZZZ
inline Jac xyzz_normalize(const Xyzz &p) {
    if (fp_is_zero(p.zz)) return jac_id();
    const Fp t = fp_inv(fp_mul(p.zz, p.zzz));
    return Jac{fp_mul(p.x, fp_mul(p.zzz, t)), fp_mul(p.y, fp_mul(p.zz, t)), fp_one()};
}

// MSM tail: the window sums are S_w = Q[w][nb] + 2^l1 sum_{b < nb} 2^b Q[w][b]
// and the MSM is sum_w 2^(c w) S_w, so every masked sum Q[w][b] carries the
// exponent c w + l1 + b (b < nb) or c w (b = nb).  One Horner pass over the
// exponents from the top down adds each term where its exponent comes up:
// c (W - 1) + l1 + nb - 1 doublings in all, instead of a doubling chain per
// window plus c per window step.  q: W * (nb + 1) XYZZ points, window-major.
inline int msm_tail_top(uint32_t c, uint32_t W, uint32_t nb, uint32_t l1) {
    return (int)(c * (W - 1) + (nb ? l1 + nb : 1)) - 1;
}
// sum over the exponents lo <= e <= hi of 2^(e - lo) T_e
inline Xyzz msm_tail_range(const Xyzz *q, uint32_t c, uint32_t W, uint32_t nb, uint32_t l1, int lo, int hi) {
    const int span = (int)(nb ? l1 + nb : 1);  // exponents a window's terms cover above c w
    Xyzz acc = xyzz_id();
    for (int e = hi; e >= lo; e--) {
        acc = xyzz_dbl(acc);
        int w = e / (int)c;
        if (w > (int)W - 1) w = (int)W - 1;
        for (; w >= 0 && e - (int)c * w < span; w--) {
            const int d = e - (int)c * w;
            const Xyzz *t = d == 0 ? &q[(size_t)w * (nb + 1) + nb]
                            : (d >= (int)l1 && d < (int)(l1 + nb)) ? &q[(size_t)w * (nb + 1) + (d - l1)]
                                                                   : nullptr;
            if (t) acc = xyzz_add(acc, *t);
        }
    }
    return acc;
}
inline Jac msm_tail(const Xyzz *q, uint32_t c, uint32_t W, uint32_t nb, uint32_t l1) {
    return xyzz_normalize(msm_tail_range(q, c, W, nb, l1, 0, msm_tail_top(c, W, nb, l1)));
}

}  // namespace zkh
