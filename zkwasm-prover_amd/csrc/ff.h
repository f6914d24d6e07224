// BN254 Fr / Fp arithmetic for gfx950: 256-bit Montgomery (R = 2^256) on
// eight 32-bit limbs.  Byte layout equals halo2's Fr([u64; 4]) / Fq([u64; 4])
// in Montgomery form, so device buffers are the caller's buffers unchanged.
//
// Multiplication (fe_mul) is Montgomery product scanning with the reduction
// folded into the same columns (FIPS), each 32x32 product accumulated by one
// v_mad_u64_u32 whose carry-out feeds v_addc_co_u32 (see mac96).  A row-wise
// CIOS variant (fe_mul_cios) is kept for A/B measurement.  All results are
// fully reduced (< modulus), which is what makes every output bit-comparable
// with the CPU restatement.
//
// Reference semantics: cuda/zprize/ff_dispatch_st.cuh:192-217 (add/sub),
// :428-529 (mul, to/from Montgomery), :655-686 (inverse, 0 -> 0),
// cuda/zprize_ff_wrapper.cuh:13-29 (small constants as v*R).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "inv_bingcd.h"

namespace zk {

struct FrParams {
    static constexpr uint32_t M[8] = {0xf0000001u, 0x43e1f593u, 0x79b97091u, 0x2833e848u,
                                      0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
    static constexpr uint32_t ONE[8] = {0x4ffffffbu, 0xac96341cu, 0x9f60cd29u, 0x36fc7695u,
                                        0x7879462eu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u};
    static constexpr uint32_t R2[8] = {0xae216da7u, 0x1bb8e645u, 0xe35c59e3u, 0x53fe3ab1u,
                                       0x53bb8085u, 0x8c49833du, 0x7f4e44a5u, 0x0216d0b1u};
    static constexpr uint32_t R3[8] = {0xb4bf0040u, 0x5e94d8e1u, 0x1cfbb6b8u, 0x2a489cbeu,
                                       0xa19fcfedu, 0x893cc664u, 0x7fcc657cu, 0x0cf8594bu};
    static constexpr uint32_t INV = 0xefffffffu;  // -r^-1 mod 2^32
    // 2^36 R^3 mod r: turns the binary GCD's cofactor into the Montgomery inverse
    static constexpr uint32_t INVC[8] = {0x962db482u, 0x4cfbeddbu, 0xd900a306u, 0xb4125eb3u,
                                         0x544b9a4du, 0x9d64b74cu, 0x0a6a2c2du, 0x2731b13cu};
};

struct FpParams {
    static constexpr uint32_t M[8] = {0xd87cfd47u, 0x3c208c16u, 0x6871ca8du, 0x97816a91u,
                                      0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
    static constexpr uint32_t ONE[8] = {0xc58f0d9du, 0xd35d438du, 0xf5c70b3du, 0x0a78eb28u,
                                        0x7879462cu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u};
    static constexpr uint32_t R2[8] = {0x538afa89u, 0xf32cfc5bu, 0xd44501fbu, 0xb5e71911u,
                                       0x0a417ff6u, 0x47ab1effu, 0xcab8351fu, 0x06d89f71u};
    static constexpr uint32_t R3[8] = {0xda1530dfu, 0xb1cd6dafu, 0xa7283db6u, 0x62f210e6u,
                                       0x0ada0afbu, 0xef7f0b0cu, 0x2d592544u, 0x20fd6e90u};
    static constexpr uint32_t INV = 0xe4866389u;  // -p^-1 mod 2^32
    // 2^36 R^3 mod p (as INVC of FrParams)
    static constexpr uint32_t INVC[8] = {0xc57e1fafu, 0x12f8df1fu, 0xa75c06e1u, 0x516b8b3cu,
                                         0xe6486ed4u, 0x448dbc80u, 0x2a2a6c68u, 0x00d762b6u};
};

template <class P>
struct alignas(32) Fe {
    uint32_t v[8];
};

using Fr = Fe<FrParams>;
using Fp = Fe<FpParams>;

// ---- limb helpers --------------------------------------------------------
// One 32-bit add / subtract with carry (borrow) in and out, 0 or 1: the
// builtins become v_add_co / v_addc_co (v_sub_co / v_subb_co) chains through
// the carry mask; the same written with 64-bit sums compiled to 64-bit adds
// and moves, about three times the instructions per limb.
__device__ __forceinline__ uint32_t add_cc(uint32_t a, uint32_t b, uint32_t &carry) {
    uint32_t c;
    const uint32_t r = __builtin_addc(a, b, carry, &c);
    carry = c;
    return r;
}
__device__ __forceinline__ uint32_t sub_bb(uint32_t a, uint32_t b, uint32_t &borrow) {
    uint32_t c;
    const uint32_t r = __builtin_subc(a, b, borrow, &c);
    borrow = c;
    return r;
}

template <class P>
__device__ __forceinline__ Fe<P> fe_zero() {
    Fe<P> r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = 0;
    return r;
}
template <class P>
__device__ __forceinline__ Fe<P> fe_one() {
    Fe<P> r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = P::ONE[i];
    return r;
}
template <class P>
__device__ __forceinline__ bool fe_is_zero(const Fe<P> &a) {
    uint32_t x = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) x |= a.v[i];
    return x == 0;
}
template <class P>
__device__ __forceinline__ bool fe_eq(const Fe<P> &a, const Fe<P> &b) {
    uint32_t x = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) x |= a.v[i] ^ b.v[i];
    return x == 0;
}

// r = t - M if t >= M else t  (t < 2M)
template <class P>
__device__ __forceinline__ Fe<P> fe_reduce_once(const Fe<P> &t) {
    Fe<P> s;
    uint32_t b = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) s.v[i] = sub_bb(t.v[i], P::M[i], b);
    Fe<P> r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = b ? t.v[i] : s.v[i];
    return r;
}

template <class P>
__device__ __forceinline__ Fe<P> fe_add(const Fe<P> &a, const Fe<P> &b) {
    Fe<P> t;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) t.v[i] = add_cc(a.v[i], b.v[i], c);
    return fe_reduce_once<P>(t);
}
template <class P>
__device__ __forceinline__ Fe<P> fe_dbl(const Fe<P> &a) { return fe_add<P>(a, a); }

template <class P>
__device__ __forceinline__ Fe<P> fe_sub(const Fe<P> &a, const Fe<P> &b) {
    Fe<P> t;
    uint32_t br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) t.v[i] = sub_bb(a.v[i], b.v[i], br);
    // add M back when the subtraction borrowed (mask instead of a branch)
    uint32_t mask = 0u - br;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) t.v[i] = add_cc(t.v[i], P::M[i] & mask, c);
    return t;
}
template <class P>
__device__ __forceinline__ Fe<P> fe_neg(const Fe<P> &a) {
    return fe_sub<P>(fe_zero<P>(), a);
}

// CIOS Montgomery multiplication, 32-bit limbs.  Inputs < M, output < M.
// The inner step (uint64)a*b + t + carry fits 64 bits exactly:
// (2^32-1)^2 + 2(2^32-1) = 2^64-1.
template <class P>
__device__ __forceinline__ Fe<P> fe_mul_cios(const Fe<P> &a, const Fe<P> &b) {
    uint32_t t[9];
#pragma unroll
    for (int i = 0; i < 9; i++) t[i] = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        uint64_t c = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            c = (uint64_t)a.v[j] * b.v[i] + t[j] + (c >> 32);
            t[j] = (uint32_t)c;
        }
        uint32_t t8 = t[8] + (uint32_t)(c >> 32);  // bounded: t < 2M before the row
        uint32_t m = t[0] * P::INV;
        c = (uint64_t)m * P::M[0] + t[0];
#pragma unroll
        for (int j = 1; j < 8; j++) {
            c = (uint64_t)m * P::M[j] + t[j] + (c >> 32);
            t[j - 1] = (uint32_t)c;
        }
        c = (uint64_t)t8 + (c >> 32);
        t[7] = (uint32_t)c;
        t[8] = (uint32_t)(c >> 32);
    }
    Fe<P> r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = t[i];
    return fe_reduce_once<P>(r);
}
// (acc:t) += x * y as a 96-bit accumulator: acc is a 64-bit VGPR pair and t
// the third word.  v_mad_u64_u32 writes its 64-bit carry-out to an SGPR lane
// mask, which v_addc_co_u32 folds into t (two VALU instructions, no moves).
__device__ __forceinline__ void mac96(uint64_t &acc, uint32_t &t, uint32_t x, uint32_t y) {
    uint64_t cc;
    asm("v_mad_u64_u32 %0, %1, %3, %4, %0\n\t"
        "v_addc_co_u32_e64 %2, %1, %2, 0, %1"
        : "+v"(acc), "=&s"(cc), "+v"(t)
        : "v"(x), "v"(y));
}
// same with an SGPR second operand (the modulus words are uniform constants)
__device__ __forceinline__ void mac96s(uint64_t &acc, uint32_t &t, uint32_t x, uint32_t y) {
    uint64_t cc;
    asm("v_mad_u64_u32 %0, %1, %3, %4, %0\n\t"
        "v_addc_co_u32_e64 %2, %1, %2, 0, %1"
        : "+v"(acc), "=&s"(cc), "+v"(t)
        : "v"(x), "s"(y));
}

// Montgomery multiplication by product scanning with the reduction folded
// into the same columns (FIPS): column i accumulates a_j*b_(i-j) and
// m_j*M_(i-j) into a 96-bit accumulator; m_i is chosen so the column's low
// word cancels, then the accumulator shifts down one word.  128 multiply-adds
// with carries caught by the 96-bit accumulator replace the 64-bit-operand
// shuffling of row-wise CIOS.  Inputs < M, output < M.
template <class P>
__device__ __forceinline__ Fe<P> fe_mul_fips(const Fe<P> &a, const Fe<P> &b) {
    uint32_t m[8];
    Fe<P> r;
    uint64_t acc = 0;
    uint32_t t = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
#pragma unroll
        for (int j = 0; j < i; j++) {
            mac96(acc, t, a.v[j], b.v[i - j]);
            mac96s(acc, t, m[j], P::M[i - j]);
        }
        mac96(acc, t, a.v[i], b.v[0]);
        m[i] = (uint32_t)acc * P::INV;
        mac96s(acc, t, m[i], P::M[0]);
        acc = (acc >> 32) | ((uint64_t)t << 32);
        t = 0;
    }
#pragma unroll
    for (int i = 8; i < 16; i++) {
#pragma unroll
        for (int j = i - 7; j < 8; j++) {
            mac96(acc, t, a.v[j], b.v[i - j]);
            mac96s(acc, t, m[j], P::M[i - j]);
        }
        r.v[i - 8] = (uint32_t)acc;
        acc = (acc >> 32) | ((uint64_t)t << 32);
        t = 0;
    }
    // the result is < 2M < 2^255, so the word above r.v[7] is zero
    return fe_reduce_once<P>(r);
}

template <class P>
__device__ __forceinline__ Fe<P> fe_mul(const Fe<P> &a, const Fe<P> &b) {
    return fe_mul_fips<P>(a, b);
}

template <class P>
__device__ __forceinline__ Fe<P> fe_sqr(const Fe<P> &a) { return fe_mul<P>(a, a); }

template <class P>
__device__ __forceinline__ Fe<P> fe_from_raw(const uint32_t *c) {
    Fe<P> r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = c[i];
    return r;
}
// canonical -> Montgomery
template <class P>
__device__ __forceinline__ Fe<P> fe_to_mont(const Fe<P> &a) {
    return fe_mul<P>(a, fe_from_raw<P>(P::R2));
}
// Montgomery -> canonical
template <class P>
__device__ __forceinline__ Fe<P> fe_from_mont(const Fe<P> &a) {
    Fe<P> one = fe_zero<P>();
    one.v[0] = 1;
    return fe_mul<P>(a, one);
}

// Binary extended Euclid on raw limbs: returns x with a*x = 1 mod M for
// a in [1, M); 0 for a = 0.  Used only on short serial tails (one inverse
// per launch or per batch), where it is far lower-latency than a^(M-2).
template <class P>
__device__ __forceinline__ bool raw_is_one(const uint32_t a[8]) {
    uint32_t x = a[0] ^ 1u;
#pragma unroll
    for (int i = 1; i < 8; i++) x |= a[i];
    return x == 0;
}
template <class P>
__device__ __forceinline__ void raw_half_mod(uint32_t x[8]) {
    // x = x/2 mod M  (x < M, M odd)
    uint32_t mask = 0u - (x[0] & 1u);
    uint32_t c = 0;
    uint32_t t[8];
#pragma unroll
    for (int i = 0; i < 8; i++) t[i] = add_cc(x[i], P::M[i] & mask, c);
    // t < 2M < 2^255: the shift never loses a carry bit
#pragma unroll
    for (int i = 0; i < 7; i++) x[i] = (t[i] >> 1) | (t[i + 1] << 31);
    x[7] = t[7] >> 1;
}
template <class P>
__device__ Fe<P> fe_inv_raw(const Fe<P> &a) {
    if (fe_is_zero<P>(a)) return fe_zero<P>();
    uint32_t u[8], v[8], x1[8], x2[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        u[i] = a.v[i];
        v[i] = P::M[i];
        x1[i] = 0;
        x2[i] = 0;
    }
    x1[0] = 1;
    // For a in [1, M) and M prime this ends in < 2*256 rounds; the bound and
    // the zero checks only guard against non-canonical input.
    for (int round = 0; round < 1024 && !raw_is_one<P>(u) && !raw_is_one<P>(v); round++) {
        uint32_t uz = 0, vz = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) { uz |= u[i]; vz |= v[i]; }
        if (!uz || !vz) break;
        while ((u[0] & 1u) == 0) {
#pragma unroll
            for (int i = 0; i < 7; i++) u[i] = (u[i] >> 1) | (u[i + 1] << 31);
            u[7] >>= 1;
            raw_half_mod<P>(x1);
        }
        while ((v[0] & 1u) == 0) {
#pragma unroll
            for (int i = 0; i < 7; i++) v[i] = (v[i] >> 1) | (v[i + 1] << 31);
            v[7] >>= 1;
            raw_half_mod<P>(x2);
        }
        // compare u >= v
        uint32_t br = 0, d[8];
#pragma unroll
        for (int i = 0; i < 8; i++) d[i] = sub_bb(u[i], v[i], br);
        if (!br) {
#pragma unroll
            for (int i = 0; i < 8; i++) u[i] = d[i];
            Fe<P> t = fe_sub<P>(fe_from_raw<P>(x1), fe_from_raw<P>(x2));
#pragma unroll
            for (int i = 0; i < 8; i++) x1[i] = t.v[i];
        } else {
            uint32_t b2 = 0;
#pragma unroll
            for (int i = 0; i < 8; i++) v[i] = sub_bb(v[i], u[i], b2);
            Fe<P> t = fe_sub<P>(fe_from_raw<P>(x2), fe_from_raw<P>(x1));
#pragma unroll
            for (int i = 0; i < 8; i++) x2[i] = t.v[i];
        }
    }
    return raw_is_one<P>(u) ? fe_from_raw<P>(x1) : fe_from_raw<P>(x2);
}
// The optimized binary GCD (inv_bingcd.h) as one out-of-line device
// function per translation unit: it is large, and inversions are rare.
__device__ __noinline__ inline bool bingcd_dev(uint32_t v[8], const uint32_t y[8], const uint32_t M[8],
                                               uint32_t minv) {
    return zkinv::bingcd(v, y, M, minv);
}

// Montgomery-domain inverse, a^-1 R for a = aR: the binary GCD's cofactor v
// (y^-1 = v 2^36 for y = aR) times INVC = 2^36 R^3 under one Montgomery
// product.  The result is checked (a * a^-1 = R); should the check fail, the
// one-bit-at-a-time extended Euclid gives (aR)^-1 and R^3 makes it a^-1 R.
// 0 -> 0.
template <class P>
__device__ Fe<P> fe_inv(const Fe<P> &a) {
    if (fe_is_zero<P>(a)) return fe_zero<P>();
    uint32_t v[8];
    const bool ok = bingcd_dev(v, a.v, P::M, P::INV);
    Fe<P> r = fe_mul<P>(fe_from_raw<P>(v), fe_from_raw<P>(P::INVC));
    if (ok && fe_eq<P>(fe_mul<P>(a, r), fe_one<P>())) return r;
    Fe<P> x = fe_inv_raw<P>(a);
    return fe_mul<P>(x, fe_from_raw<P>(P::R3));
}

// small constant v as v*R (cuda/zprize_ff_wrapper.cuh:13-29)
template <class P>
__device__ __forceinline__ Fe<P> fe_from_u32(uint32_t v) {
    Fe<P> c = fe_zero<P>();
    c.v[0] = v;
    return fe_to_mont<P>(c);
}

// vectorised global loads/stores (2 x dwordx4)
template <class P>
__device__ __forceinline__ Fe<P> fe_load(const Fe<P> *p) {
    const uint4 *q = reinterpret_cast<const uint4 *>(p);
    uint4 a = q[0], b = q[1];
    Fe<P> r;
    r.v[0] = a.x; r.v[1] = a.y; r.v[2] = a.z; r.v[3] = a.w;
    r.v[4] = b.x; r.v[5] = b.y; r.v[6] = b.z; r.v[7] = b.w;
    return r;
}
template <class P>
__device__ __forceinline__ void fe_store(Fe<P> *p, const Fe<P> &r) {
    uint4 *q = reinterpret_cast<uint4 *>(p);
    q[0] = make_uint4(r.v[0], r.v[1], r.v[2], r.v[3]);
    q[1] = make_uint4(r.v[4], r.v[5], r.v[6], r.v[7]);
}

// operator sugar for kernels written as formulas
template <class P> __device__ __forceinline__ Fe<P> operator+(const Fe<P> &a, const Fe<P> &b) { return fe_add<P>(a, b); }
template <class P> __device__ __forceinline__ Fe<P> operator-(const Fe<P> &a, const Fe<P> &b) { return fe_sub<P>(a, b); }
template <class P> __device__ __forceinline__ Fe<P> operator*(const Fe<P> &a, const Fe<P> &b) { return fe_mul<P>(a, b); }

}  // namespace zk
