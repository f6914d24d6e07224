// Host-side BN254 Fr arithmetic (Montgomery form, R = 2^256, four 64-bit
// limbs: the bytes of halo2's Fr and of the device Fr).  Used for the few
// scalar computations of the opening step (Lagrange interpolation of at most
// a handful of points, products over the point set); everything of size n
// runs on the device.
#pragma once
#include <stdint.h>
#include <string.h>

namespace zkh {

struct Fr {
    uint64_t l[4];
};

constexpr uint64_t kR[4] = {0x43e1f593f0000001ull, 0x2833e84879b97091ull, 0xb85045b68181585dull,
                            0x30644e72e131a029ull};
constexpr uint64_t kRInv = 0xc2e1f593efffffffull;  // -r^-1 mod 2^64
constexpr uint64_t kR2[4] = {0x1bb8e645ae216da7ull, 0x53fe3ab1e35c59e3ull, 0x8c49833d53bb8085ull,
                             0x0216d0b17f4e44a5ull};

inline bool fr_geq_r(const uint64_t a[4]) {
    for (int i = 3; i >= 0; i--) {
        if (a[i] != kR[i]) return a[i] > kR[i];
    }
    return true;
}
inline void fr_sub_r(uint64_t a[4]) {
    unsigned __int128 b = 0;
    for (int i = 0; i < 4; i++) {
        unsigned __int128 d = (unsigned __int128)a[i] - kR[i] - (uint64_t)b;
        a[i] = (uint64_t)d;
        b = (d >> 64) & 1;
    }
}
inline Fr fr_zero() { return Fr{{0, 0, 0, 0}}; }
inline bool fr_is_zero(const Fr &a) { return !(a.l[0] | a.l[1] | a.l[2] | a.l[3]); }
inline bool fr_eq(const Fr &a, const Fr &b) { return memcmp(&a, &b, sizeof(Fr)) == 0; }

inline Fr fr_add(const Fr &a, const Fr &b) {
    Fr r;
    unsigned __int128 c = 0;
    for (int i = 0; i < 4; i++) {
        c += (unsigned __int128)a.l[i] + b.l[i];
        r.l[i] = (uint64_t)c;
        c >>= 64;
    }
    if (c || fr_geq_r(r.l)) fr_sub_r(r.l);  // a + b < 2r < 2^256: c is 0
    return r;
}
inline Fr fr_sub(const Fr &a, const Fr &b) {
    Fr r;
    unsigned __int128 br = 0;
    for (int i = 0; i < 4; i++) {
        unsigned __int128 d = (unsigned __int128)a.l[i] - b.l[i] - (uint64_t)br;
        r.l[i] = (uint64_t)d;
        br = (d >> 64) & 1;
    }
    if (br) {  // add r back
        unsigned __int128 c = 0;
        for (int i = 0; i < 4; i++) {
            c += (unsigned __int128)r.l[i] + kR[i];
            r.l[i] = (uint64_t)c;
            c >>= 64;
        }
    }
    return r;
}
// CIOS Montgomery product a b / 2^256 mod r
inline Fr fr_mul(const Fr &a, const Fr &b) {
    uint64_t t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; i++) {
        unsigned __int128 c = 0;
        for (int j = 0; j < 4; j++) {
            c += (unsigned __int128)a.l[j] * b.l[i] + t[j];
            t[j] = (uint64_t)c;
            c >>= 64;
        }
        c += t[4];
        t[4] = (uint64_t)c;
        t[5] = (uint64_t)(c >> 64);
        const uint64_t m = t[0] * kRInv;
        c = (unsigned __int128)m * kR[0] + t[0];
        c >>= 64;
        for (int j = 1; j < 4; j++) {
            c += (unsigned __int128)m * kR[j] + t[j];
            t[j - 1] = (uint64_t)c;
            c >>= 64;
        }
        c += t[4];
        t[3] = (uint64_t)c;
        t[4] = t[5] + (uint64_t)(c >> 64);
    }
    Fr r{{t[0], t[1], t[2], t[3]}};
    if (t[4] || fr_geq_r(r.l)) fr_sub_r(r.l);
    return r;
}
inline Fr fr_from_u64(uint64_t v) { return fr_mul(Fr{{v, 0, 0, 0}}, Fr{{kR2[0], kR2[1], kR2[2], kR2[3]}}); }
inline Fr fr_one() { return fr_from_u64(1); }
inline Fr fr_neg(const Fr &a) { return fr_sub(fr_zero(), a); }
// a^(r-2): 0 maps to 0
inline Fr fr_inv(const Fr &a) {
    uint64_t e[4] = {kR[0] - 2, kR[1], kR[2], kR[3]};
    Fr r = fr_one();
    for (int i = 3; i >= 0; i--)
        for (int b = 63; b >= 0; b--) {
            r = fr_mul(r, r);
            if ((e[i] >> b) & 1) r = fr_mul(r, a);
        }
    return r;
}

}  // namespace zkh
