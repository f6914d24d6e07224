// BN254 Fp in nine 29-bit limbs with lazy Montgomery reduction (R' = 2^261),
// for the bucket accumulation, where field multiplication is the bound.
//
// Why: with 32-bit limbs every 32x32 multiply-add overflows the 64-bit
// accumulator often enough that each one needs a carry-catching add (128
// multiply-adds + 128 carry adds per product).  With 29-bit limbs a partial
// product is < 2^58, so a whole column of the product-scanning Montgomery
// multiplication (at most 18 terms) sums in a 64-bit register with no carries:
// 162 multiply-adds, one 29-bit shift per column, nothing else.  R' = 2^261 is
// far above 4p, so inputs may stay unreduced (< 2p after add / sub, < 32p for a
// coordinate converted from the 32-bit form) and the product is < 2p without a
// final subtraction.
//
// Values enter from the library's 32-bit Montgomery form (x * 2^256 mod p)
// by f29_from32, which only shifts the 256-bit value left by 5 (x * 2^261, not
// reduced: < 32p), and leave by f29_to32 (canonical, < p).
#pragma once
#include "ec.h"
#include "ff.h"

namespace zk {

struct F29 {
    uint32_t l[9];
};

constexpr uint32_t kM29 = (1u << 29) - 1;
// p in 29-bit limbs, 2p, -p^-1 mod 2^29, -p^-1 mod 2^5, 2^261 mod p
constexpr uint32_t kP29[9] = {0x187cfd47u, 0x010460b6u, 0x1c72a34fu, 0x02d522d0u, 0x1585d978u,
                              0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
constexpr uint32_t kP2x29[9] = {0x10f9fa8eu, 0x0208c16du, 0x18e5469eu, 0x05aa45a1u, 0x0b0bb2f0u,
                                0x05b68181u, 0x014dc282u, 0x1cb84c68u, 0x0060c89cu};
constexpr uint32_t kP4x29[9] = {0x01f3f51cu, 0x041182dbu, 0x11ca8d3cu, 0x0b548b43u, 0x161765e0u,
                                0x0b6d0302u, 0x029b8504u, 0x197098d0u, 0x00c19139u};
constexpr uint32_t kP6x29[9] = {0x12edefaau, 0x061a4448u, 0x0aafd3dau, 0x10fed0e5u, 0x012318d0u,
                                0x11238484u, 0x03e94786u, 0x1628e538u, 0x012259d6u};
constexpr uint32_t kP8x29[9] = {0x03e7ea38u, 0x082305b6u, 0x03951a78u, 0x16a91687u, 0x0c2ecbc0u,
                                0x16da0605u, 0x05370a08u, 0x12e131a0u, 0x01832273u};
constexpr uint32_t kPInv29 = 0x04866389u;
constexpr uint32_t kPInv5 = 9u;
constexpr uint32_t kOne29[9] = {0x157ccc21u, 0x141c2758u, 0x185230d3u, 0x014c0419u, 0x0aa36fb9u,
                                0x1d4240ceu, 0x11d54c07u, 0x052ac7a8u, 0x000dc836u};

__device__ __forceinline__ F29 f29_const(const uint32_t (&c)[9]) {
    F29 r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.l[i] = c[i];
    return r;
}
__device__ __forceinline__ F29 f29_one() { return f29_const(kOne29); }

// a * b * 2^-261 mod p, < 2p, for a * b < 2^261 * p (e.g. a < 32p, b < 2p)
__device__ __forceinline__ F29 f29_mul_chain(const F29 &a, const F29 &b) {
    uint32_t m[9];
    F29 r;
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < 17; k++) {
        const int lo = k < 9 ? 0 : k - 8, hi = k < 9 ? k : 8;
#pragma unroll
        for (int i = lo; i <= hi; i++) acc += (uint64_t)a.l[i] * b.l[k - i];
#pragma unroll
        for (int i = lo; i <= hi; i++)
            if (i < k) acc += (uint64_t)m[i] * kP29[k - i];
        if (k < 9) {
            m[k] = ((uint32_t)acc * kPInv29) & kM29;
            acc += (uint64_t)m[k] * kP29[0];
        } else {
            r.l[k - 9] = (uint32_t)acc & kM29;
        }
        acc >>= 29;
    }
    r.l[8] = (uint32_t)acc;
    return r;
}

// The modulus of a 29-bit product: Fp (every curve kernel) or Fr (the NTT's
// twiddle products).  Both are < 2^254, so every bound stated for p holds for r.
struct Fp29Mod {
    static constexpr uint32_t P[9] = {0x187cfd47u, 0x010460b6u, 0x1c72a34fu, 0x02d522d0u, 0x1585d978u,
                                      0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
    static constexpr uint32_t PINV = kPInv29;
};
struct Fr29Mod {
    static constexpr uint32_t P[9] = {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u,
                                      0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
    static constexpr uint32_t PINV = 0x0fffffffu;  // -r^-1 mod 2^29
};

// the same product with each column's a*b terms (two chains) and m*p terms
// summed independently: shorter dependency chains of multiply-adds (+6%
// measured); this is the one the kernels use
template <class Q = Fp29Mod>
__device__ __forceinline__ F29 f29_mul(const F29 &a, const F29 &b) {
    uint32_t m[9];
    F29 r;
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < 17; k++) {
        const int lo = k < 9 ? 0 : k - 8, hi = k < 9 ? k : 8;
        uint64_t t0 = 0, t1 = 0;
#pragma unroll
        for (int i = lo; i <= hi; i++) {
            if ((i - lo) & 1)
                t1 += (uint64_t)a.l[i] * b.l[k - i];
            else
                t0 += (uint64_t)a.l[i] * b.l[k - i];
        }
        uint64_t u = 0;
#pragma unroll
        for (int i = lo; i <= hi; i++)
            if (i < k) u += (uint64_t)m[i] * Q::P[k - i];
        acc += t0 + t1 + u;
        if (k < 9) {
            m[k] = ((uint32_t)acc * Q::PINV) & kM29;
            acc += (uint64_t)m[k] * Q::P[0];
        } else {
            r.l[k - 9] = (uint32_t)acc & kM29;
        }
        acc >>= 29;
    }
    r.l[8] = (uint32_t)acc;
    return r;
}

// (a * b + c * d) * 2^-261 mod p, < 2p, with one reduction for both
// products (the sum of the two a*b-type terms and the m*p terms of a column
// is at most 27 values < 2^58); valid for a * b + c * d < 2^261 * p, e.g. all
// four < 2p
__device__ __forceinline__ F29 f29_mul2(const F29 &a, const F29 &b, const F29 &c, const F29 &d) {
    uint32_t m[9];
    F29 r;
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < 17; k++) {
        const int lo = k < 9 ? 0 : k - 8, hi = k < 9 ? k : 8;
        uint64_t t0 = 0, t1 = 0;
#pragma unroll
        for (int i = lo; i <= hi; i++) {
            t0 += (uint64_t)a.l[i] * b.l[k - i];
            t1 += (uint64_t)c.l[i] * d.l[k - i];
        }
        uint64_t u = 0;
#pragma unroll
        for (int i = lo; i <= hi; i++)
            if (i < k) u += (uint64_t)m[i] * kP29[k - i];
        acc += t0 + t1 + u;
        if (k < 9) {
            m[k] = ((uint32_t)acc * kPInv29) & kM29;
            acc += (uint64_t)m[k] * kP29[0];
        } else {
            r.l[k - 9] = (uint32_t)acc & kM29;
        }
        acc >>= 29;
    }
    r.l[8] = (uint32_t)acc;
    return r;
}

// a^2, cross products once and doubled
__device__ __forceinline__ F29 f29_sqr(const F29 &a) {
    uint32_t m[9];
    F29 r;
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < 17; k++) {
        const int lo = k < 9 ? 0 : k - 8, hi = k < 9 ? k : 8;
        uint64_t cross = 0;
#pragma unroll
        for (int i = lo; i <= hi; i++)
            if (i < k - i) cross += (uint64_t)a.l[i] * a.l[k - i];
        acc += cross + cross;
        if ((k & 1) == 0 && k / 2 <= 8) acc += (uint64_t)a.l[k / 2] * a.l[k / 2];
#pragma unroll
        for (int i = lo; i <= hi; i++)
            if (i < k) acc += (uint64_t)m[i] * kP29[k - i];
        if (k < 9) {
            m[k] = ((uint32_t)acc * kPInv29) & kM29;
            acc += (uint64_t)m[k] * kP29[0];
        } else {
            r.l[k - 9] = (uint32_t)acc & kM29;
        }
        acc >>= 29;
    }
    r.l[8] = (uint32_t)acc;
    return r;
}

// r - 2p if that is >= 0 (r normalised, < 4p), else r
__device__ __forceinline__ F29 f29_sub2p_if(const F29 &r) {
    F29 d;
    int32_t br = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        int32_t s = (int32_t)r.l[i] - (int32_t)kP2x29[i] + br;
        d.l[i] = (uint32_t)s & kM29;
        br = s >> 29;
    }
    return br < 0 ? r : d;
}

// a + b for a, b < 2p: < 2p
__device__ __forceinline__ F29 f29_add(const F29 &a, const F29 &b) {
    F29 r;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        uint32_t s = a.l[i] + b.l[i] + c;
        r.l[i] = s & kM29;
        c = s >> 29;
    }
    return f29_sub2p_if(r);
}

// a - b for a, b < 2p: a + 2p - b, then < 2p
__device__ __forceinline__ F29 f29_sub(const F29 &a, const F29 &b) {
    F29 r;
    int32_t br = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        int32_t s = (int32_t)a.l[i] - (int32_t)b.l[i] + (int32_t)kP2x29[i] + br;
        r.l[i] = (uint32_t)s & kM29;
        br = s >> 29;
    }
    return f29_sub2p_if(r);
}

__device__ __forceinline__ F29 f29_dbl(const F29 &a) { return f29_add(a, a); }

// Unreduced forms for values that only feed products or one final
// reduction: a + b, and a - b + K p for K = 2, 4, 8 (b < K p + a).  A product
// stays < 2p while a * b < 2^261 p (p < 2^253.6: e.g. 10p x 10p, 32p x 2p),
// and a * b + c * d likewise; normalised limbs (< 2^29) keep every product
// column below 2^63 whatever the value.
__device__ __forceinline__ F29 f29_add_nr(const F29 &a, const F29 &b) {
    F29 r;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        uint32_t s = a.l[i] + b.l[i] + c;
        r.l[i] = s & kM29;
        c = s >> 29;
    }
    return r;
}
template <int K>  // K = 2, 4 or 8: adds K p
__device__ __forceinline__ F29 f29_sub_nr(const F29 &a, const F29 &b) {
    static_assert(K == 2 || K == 4 || K == 8, "K p constants exist for K = 2, 4, 8");
    F29 r;
    int32_t br = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const uint32_t kp = K == 8 ? kP8x29[i] : K == 4 ? kP4x29[i] : kP2x29[i];
        int32_t s = (int32_t)a.l[i] - (int32_t)b.l[i] + (int32_t)kp + br;
        r.l[i] = (uint32_t)s & kM29;
        br = s >> 29;
    }
    return r;
}
// a - b - 2c + 6p in one carry pass, for a, b, c < 2p: in (0, 8p)
__device__ __forceinline__ F29 f29_sub_sub2_nr(const F29 &a, const F29 &b, const F29 &c) {
    F29 r;
    int32_t br = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        int32_t s = (int32_t)a.l[i] - (int32_t)b.l[i] - (int32_t)(c.l[i] << 1) + (int32_t)kP6x29[i] + br;
        r.l[i] = (uint32_t)s & kM29;
        br = s >> 29;
    }
    return r;
}
// a < 8p -> < 2p
__device__ __forceinline__ F29 f29_reduce8(const F29 &a) {
    F29 d;
    int32_t br = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        int32_t s = (int32_t)a.l[i] - (int32_t)kP4x29[i] + br;
        d.l[i] = (uint32_t)s & kM29;
        br = s >> 29;
    }
    return f29_sub2p_if(br < 0 ? a : d);
}

// a == 0 mod p for a < 2p
__device__ __forceinline__ bool f29_is_zero(const F29 &a) {
    uint32_t z = 0, q = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        z |= a.l[i];
        q |= a.l[i] ^ kP29[i];
    }
    return z == 0 || q == 0;
}

// 32-bit Montgomery (x 2^256, < p) -> x 2^261 (the value shifted left by 5;
// < 32p, not reduced)
template <class P>
__device__ __forceinline__ F29 f29_from32(const Fe<P> &x) {
    F29 r;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const int bit = 29 * i - 5;  // bit of x that lands at limb i's bit 0
        uint32_t v;
        if (bit < 0) {
            v = x.v[0] << 5;
        } else {
            const int w = bit >> 5, o = bit & 31;
            v = x.v[w] >> o;
            if (o && w + 1 < 8) v |= x.v[w + 1] << (32 - o);
        }
        r.l[i] = v & kM29;
    }
    return r;
}

// x 2^261 (< 2p) -> canonical x 2^256 mod p in 32-bit limbs
__device__ __forceinline__ Fp f29_to32(const F29 &a) {
    // reduce below p
    F29 v;
    {
        int32_t br = 0;
        F29 d;
#pragma unroll
        for (int i = 0; i < 9; i++) {
            int32_t s = (int32_t)a.l[i] - (int32_t)kP29[i] + br;
            d.l[i] = (uint32_t)s & kM29;
            br = s >> 29;
        }
        v = br < 0 ? a : d;
    }
    // divide by 32 mod p: add k p with k = -v p^-1 mod 32, then shift right by 5
    const uint32_t k = (v.l[0] * kPInv5) & 31u;
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        c += (uint64_t)v.l[i] + (uint64_t)k * kP29[i];
        v.l[i] = (uint32_t)c & kM29;
        c >>= 29;
    }
    // 256-bit value = v >> 5, repacked into 32-bit words
    Fp r;
#pragma unroll
    for (int w = 0; w < 8; w++) {
        const int bit = 32 * w + 5;  // bit of v that lands at word w's bit 0
        const int i = bit / 29, o = bit % 29;
        uint32_t x = v.l[i] >> o;
        if (i + 1 < 9) x |= v.l[i + 1] << (29 - o);
        if (o > 26 && i + 2 < 9) x |= v.l[i + 2] << (58 - o);
        r.v[w] = x;
    }
    return r;
}

// 32-bit limbs (any value < 2^256) as nine 29-bit limbs of the same value
__device__ __forceinline__ F29 f29_repack(const uint32_t x[8]) {
    F29 r;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const int bit = 29 * i, w = bit >> 5, o = bit & 31;
        uint32_t v = x[w] >> o;
        if (o > 3 && w + 1 < 8) v |= x[w + 1] << (32 - o);
        r.l[i] = v & kM29;
    }
    return r;
}
// nine normalised 29-bit limbs of a value < 2^256 as eight 32-bit words
__device__ __forceinline__ void f29_unpack(const F29 &a, uint32_t x[8]) {
#pragma unroll
    for (int w = 0; w < 8; w++) {
        const int bit = 32 * w, i = bit / 29, o = bit % 29;
        uint32_t v = a.l[i] >> o;
        if (i + 1 < 9) v |= a.l[i + 1] << (29 - o);
        if (o > 26 && i + 2 < 9) v |= a.l[i + 2] << (58 - o);
        x[w] = v;
    }
}
// a * w in Fr's 32-bit Montgomery form for a < r and w29 = f29_from32(w)
// (w 2^261 mod r up to 32 r): the 29-bit product a w29 2^-261 is a w 2^256
// mod r itself, < 2r, so repacking both ways and one conditional
// subtraction replace the 32-bit product (about 270 VALU instructions
// against 318)
__device__ __forceinline__ Fr fr_mul_w29(const Fr &a, const F29 &w29) {
    const F29 t = f29_mul<Fr29Mod>(f29_repack(a.v), w29);
    Fr r;
    f29_unpack(t, r.v);
    return fe_reduce_once<FrParams>(r);
}

}  // namespace zk

namespace zk {

// XYZZ point over F29 (x = X/ZZ, y = Y/ZZZ); a separate flag marks the identity
struct X29 {
    F29 x, y, zz, zzz;
};

// dbl-2008-s-1 (a = 0)
__device__ __forceinline__ X29 x29_dbl(const X29 &p) {
    F29 U = f29_dbl(p.y);
    F29 V = f29_sqr(U);
    F29 W = f29_mul(U, V);
    F29 S = f29_mul(p.x, V);
    F29 X2 = f29_sqr(p.x);
    F29 M = f29_add(f29_dbl(X2), X2);
    X29 r;
    r.x = f29_sub(f29_sqr(M), f29_dbl(S));
    r.y = f29_sub(f29_mul(M, f29_sub(S, r.x)), f29_mul(W, p.y));
    r.zz = f29_mul(V, p.zz);
    r.zzz = f29_mul(W, p.zzz);
    return r;
}

// The empty accumulator: zz = zzz = 1, so that the first two products of the
// mixed addition below return the new point's coordinates reduced.
__device__ __forceinline__ void x29_clear(X29 &p, bool &empty) {
    p.zz = f29_one();
    p.zzz = p.zz;
    empty = true;
}

// acc += (x2, y2) (madd-2008-s), x2, y2 < 32p used only as multiplication
// operands.  An empty acc takes the point itself (U2 = x2 * 1, S2 = y2 * 1 are
// its reduced coordinates), so lanes starting a bucket run do no extra work.
// The accumulator's x stays unreduced (< 8p, normalised limbs) between
// additions: it only enters products and P = U2 - X1 + 8p (< 10p); X3 =
// R^2 - PPP - 2Q + 6p comes out of one carry pass.  Store it with
// x29_store_acc, which reduces x below 2p as every other X29 consumer expects.
__device__ __forceinline__ void x29_acc_aff(X29 &p, bool &empty, const F29 &x2, const F29 &y2) {
    F29 U2 = f29_mul(x2, p.zz);
    F29 S2 = f29_mul(y2, p.zzz);
    if (empty) {
        p.x = U2;
        p.y = S2;
        empty = false;
        return;
    }
    const F29 P = f29_sub_nr<8>(U2, p.x);  // < 10p: only multiplied
    const F29 R = f29_sub_nr<2>(S2, p.y);  // < 4p: only multiplied
    const F29 PP = f29_sqr(P);             // < 2p; zero exactly when P is
    if (f29_is_zero(PP)) {
        if (f29_is_zero(f29_sub2p_if(R))) {
            p = x29_dbl(p);  // the point equals the accumulator
        } else {
            x29_clear(p, empty);
        }
        return;
    }
    const F29 PPP = f29_mul(P, PP);
    const F29 Q = f29_mul(p.x, PP);
    const F29 X3 = f29_sub_sub2_nr(f29_sqr(R), PPP, Q);  // < 8p
    // R (Q - X3) - Y1 PPP as R (Q - X3 + 8p) + (2p - Y1) PPP: one reduction,
    // operands < 4p, < 10p, <= 2p, < 2p (sum of products < 44 p^2)
    const F29 zero = {{0, 0, 0, 0, 0, 0, 0, 0, 0}};
    F29 Y3 = f29_mul2(R, f29_sub_nr<8>(Q, X3), f29_sub_nr<2>(zero, p.y), PPP);
    p.zz = f29_mul(p.zz, PP);
    p.zzz = f29_mul(p.zzz, PPP);
    p.x = X3;
    p.y = Y3;
}

// p += q for two XYZZ points (add-2008-s); the flags mark the identity.
// Every coordinate is < 2p, as F29 results are.
__device__ __forceinline__ void x29_add(X29 &p, bool &pe, const X29 &q, bool qe) {
    if (qe) return;
    if (pe) {
        p = q;
        pe = false;
        return;
    }
    F29 U1 = f29_mul(p.x, q.zz);
    F29 U2 = f29_mul(q.x, p.zz);
    F29 S1 = f29_mul(p.y, q.zzz);
    F29 S2 = f29_mul(q.y, p.zzz);
    F29 P = f29_sub(U2, U1);
    F29 R = f29_sub(S2, S1);
    if (f29_is_zero(P)) {
        if (f29_is_zero(R))
            p = x29_dbl(p);
        else
            x29_clear(p, pe);
        return;
    }
    F29 PP = f29_sqr(P);
    F29 PPP = f29_mul(P, PP);
    F29 Q = f29_mul(U1, PP);
    F29 X3 = f29_sub(f29_sub(f29_sqr(R), PPP), f29_dbl(Q));
    const F29 zero = {{0, 0, 0, 0, 0, 0, 0, 0, 0}};
    F29 Y3 = f29_mul2(R, f29_sub(Q, X3), f29_sub(zero, S1), PPP);
    p.zz = f29_mul(f29_mul(p.zz, q.zz), PP);
    p.zzz = f29_mul(f29_mul(p.zzz, q.zzz), PPP);
    p.x = X3;
    p.y = Y3;
}

// Storage form of an X29 point between kernels: the 36 limbs, 144 B,
// loaded and stored as nine 16-byte words; zz = 0 marks the identity.
struct alignas(16) X29S {
    uint4 q[9];
};
__device__ __forceinline__ void x29_store(X29S *d, const X29 &p, bool empty) {
    uint32_t w[36];
#pragma unroll
    for (int i = 0; i < 9; i++) {
        w[i] = p.x.l[i];
        w[9 + i] = p.y.l[i];
        w[18 + i] = empty ? 0u : p.zz.l[i];
        w[27 + i] = p.zzz.l[i];
    }
#pragma unroll
    for (int k = 0; k < 9; k++) d->q[k] = make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]);
}
// x29_store for an accumulator of x29_acc_aff (x < 8p): x reduced first
__device__ __forceinline__ void x29_store_acc(X29S *d, X29 p, bool empty) {
    p.x = f29_reduce8(p.x);
    x29_store(d, p, empty);
}
__device__ __forceinline__ X29 x29_load(const X29S *s, bool &empty) {
    uint32_t w[36];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const uint4 v = s->q[k];
        w[4 * k] = v.x;
        w[4 * k + 1] = v.y;
        w[4 * k + 2] = v.z;
        w[4 * k + 3] = v.w;
    }
    X29 p;
    uint32_t z = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        p.x.l[i] = w[i];
        p.y.l[i] = w[9 + i];
        p.zz.l[i] = w[18 + i];
        p.zzz.l[i] = w[27 + i];
        z |= w[18 + i];
    }
    empty = z == 0;
    return p;
}

// back to the library's XYZZ (canonical 32-bit Montgomery coordinates)
__device__ __forceinline__ G1Xyzz x29_to_xyzz(const X29 &p, bool empty) {
    if (empty) return xyzz_id();
    G1Xyzz r;
    r.x = f29_to32(p.x);
    r.y = f29_to32(p.y);
    r.zz = f29_to32(p.zz);
    r.zzz = f29_to32(p.zzz);
    return r;
}

}  // namespace zk
