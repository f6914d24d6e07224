// Modular inversion by the optimized binary GCD of T. Pornin ("Optimized
// Binary GCD for Modular Inversion", 2020): the binary extended Euclid run kBits
// steps at a time on 64-bit approximations of (a, b) -- the low kBits bits
// exact, the top 64 - kBits bits of the larger operand's length -- which
// yields a 2x2 matrix (f0 g0; f1 g1) with |entries| <= 2^kBits; the matrix is
// then applied once to the full 256-bit a, b (exact division by 2^kBits, signs
// fixed by negation) and to the cofactors u, v modulo M, the latter with one
// Montgomery word reduction (a factor 2^-32 per step, folded into the final
// constant).  Invariant: a = u y 2^(-kBits t) 2^(32 t), b likewise with v
// (mod M) after t steps, so when a reaches 0 and b = gcd = 1,
// y^-1 = v 2^((32 - kBits) T) with T the number of steps run.
//
// kBits = 30 keeps every signed 32 x 31-bit product sum below 2^63.  kSteps =
// 18 steps give 540 binary iterations, above the 2 * 254 - 1 the algorithm
// needs for a 254-bit modulus; extra steps leave (a, b) = (0, 1) unchanged.
// Callers verify the result with one product and keep a slower exact path,
// so correctness does not rest on the approximation bound.
//
// Plain integer code on eight 32-bit limbs, usable on the host and the device.
#pragma once
#include <stdint.h>

#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)
#define ZK_HD __host__ __device__ __forceinline__
#else
#define ZK_HD inline
#endif

namespace zkinv {

constexpr int kBits = 30;
constexpr int kSteps = 18;

ZK_HD int bitlen8(const uint32_t x[8]) {
    for (int i = 7; i >= 0; i--)
        if (x[i]) return 32 * i + 32 - __builtin_clz(x[i]);
    return 0;
}

// 64-bit window of x starting at bit s (bits past the top read as zero)
ZK_HD uint64_t window64(const uint32_t x[8], int s) {
    const int w = s >> 5, o = s & 31;
    uint64_t lo = 0, mid = 0, hi = 0;
    if (w < 8) lo = x[w];
    if (w + 1 < 8) mid = x[w + 1];
    if (w + 2 < 8) hi = x[w + 2];
    const uint64_t v = lo | (mid << 32);
    return o ? (v >> o) | (hi << (64 - o)) : v;
}

// x <- (a f + b g) >> kBits as a signed value; returns true when the result
// was negative (x then holds its absolute value)
ZK_HD bool apply_row(uint32_t x[8], const uint32_t a[8], const uint32_t b[8], int64_t f, int64_t g) {
    uint32_t r[9];
    int64_t acc = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        acc += (int64_t)(uint64_t)a[i] * f + (int64_t)(uint64_t)b[i] * g;
        r[i] = (uint32_t)acc;
        acc >>= 32;  // arithmetic: the carry keeps its sign
    }
    r[8] = (uint32_t)acc;
    const bool neg = acc < 0;
    // shift the 288-bit two's-complement value right by kBits
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = (r[i] >> kBits) | (r[i + 1] << (32 - kBits));
    if (neg) {
        uint32_t c = 1;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint64_t t = (uint64_t)(uint32_t)~x[i] + c;
            x[i] = (uint32_t)t;
            c = (uint32_t)(t >> 32);
        }
    }
    return neg;
}

// x <- (u f + v g) 2^-32 mod M, in [0, M); M odd below 2^254, minv = -M^-1 mod 2^32
ZK_HD void apply_cofactor(uint32_t x[8], const uint32_t u[8], const uint32_t v[8], int64_t f, int64_t g,
                          const uint32_t M[8], uint32_t minv) {
    uint32_t r[9];
    int64_t acc = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        acc += (int64_t)(uint64_t)u[i] * f + (int64_t)(uint64_t)v[i] * g;
        r[i] = (uint32_t)acc;
        acc >>= 32;
    }
    // t = r + acc 2^256 (signed, |t| < 2^31 * 2M); add q M with q making the
    // low word zero, then drop that word: t' = (t + q M) / 2^32
    const uint32_t q = r[0] * minv;
    int64_t c = 0;
    uint32_t s[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int64_t t = (int64_t)(uint64_t)r[i] + (int64_t)((uint64_t)q * M[i] & 0xffffffffull) + c;
        // the product's high half joins the next word
        c = (t >> 32) + (int64_t)((uint64_t)q * M[i] >> 32);
        if (i) s[i - 1] = (uint32_t)t;
    }
    int64_t top = c + acc;  // the signed word above s[6]
    s[7] = (uint32_t)top;
    int64_t sign = top >> 32;  // 0 or -1: |t'| < 2M < 2^255
    // bring t' into [0, M): add M while negative, subtract while >= M
    for (int round = 0; round < 3; round++) {
        if (sign < 0) {
            uint32_t cc = 0;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const uint64_t t = (uint64_t)s[i] + M[i] + cc;
                s[i] = (uint32_t)t;
                cc = (uint32_t)(t >> 32);
            }
            sign += cc;
        } else {
            uint32_t d[8], br = 0;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const uint64_t t = (uint64_t)s[i] - M[i] - br;
                d[i] = (uint32_t)t;
                br = (uint32_t)(t >> 63);
            }
            if (br) break;  // s < M
#pragma unroll
            for (int i = 0; i < 8; i++) s[i] = d[i];
        }
    }
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = s[i];
}

// v with y^-1 = v 2^((32 - kBits) kSteps) mod M, for 0 < y < M; returns false
// when b did not end at 1 (y not invertible, or the bound failed)
ZK_HD bool bingcd(uint32_t v_out[8], const uint32_t y[8], const uint32_t M[8], uint32_t minv) {
    uint32_t a[8], b[8], u[8], v[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        a[i] = y[i];
        b[i] = M[i];
        u[i] = 0;
        v[i] = 0;
    }
    u[0] = 1;
    for (int step = 0; step < kSteps; step++) {
        int n = bitlen8(a);
        const int nb = bitlen8(b);
        if (nb > n) n = nb;
        if (n < 64) n = 64;
        const uint64_t lowmask = (1ull << kBits) - 1;
        uint64_t ab = (window64(a, 0) & lowmask) | (window64(a, n - (64 - kBits)) << kBits);
        uint64_t bb = (window64(b, 0) & lowmask) | (window64(b, n - (64 - kBits)) << kBits);
        int64_t f0 = 1, g0 = 0, f1 = 0, g1 = 1;
        for (int j = 0; j < kBits; j++) {
            if (ab & 1) {
                if (ab < bb) {
                    const uint64_t t = ab;
                    ab = bb;
                    bb = t;
                    int64_t tf = f0, tg = g0;
                    f0 = f1;
                    g0 = g1;
                    f1 = tf;
                    g1 = tg;
                }
                ab -= bb;
                f0 -= f1;
                g0 -= g1;
            }
            ab >>= 1;
            f1 *= 2;
            g1 *= 2;
        }
        uint32_t na[8], nbv[8], nu[8], nv[8];
        if (apply_row(na, a, b, f0, g0)) {
            f0 = -f0;
            g0 = -g0;
        }
        if (apply_row(nbv, a, b, f1, g1)) {
            f1 = -f1;
            g1 = -g1;
        }
        apply_cofactor(nu, u, v, f0, g0, M, minv);
        apply_cofactor(nv, u, v, f1, g1, M, minv);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            a[i] = na[i];
            b[i] = nbv[i];
            u[i] = nu[i];
            v[i] = nv[i];
        }
    }
    uint32_t one = b[0] ^ 1u;
#pragma unroll
    for (int i = 1; i < 8; i++) one |= b[i];
#pragma unroll
    for (int i = 0; i < 8; i++) v_out[i] = v[i];
    return one == 0;
}

}  // namespace zkinv
