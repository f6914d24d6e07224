// Fiat-Shamir transcript of the halo2 prover: Blake2b-512 with the
// personalisation "Halo2-Transcript" (halo2_proofs' Blake2bWrite, a
// dependency of the reference's create_proof, src/lib.rs:205-1194).
//   common point:  absorb 0x01, then x and y, each 32 canonical little-endian
//                  bytes (the identity cannot be absorbed)
//   common scalar: absorb 0x02, then the 32 canonical little-endian bytes
//   challenge:     absorb 0x00, then Blake2b-512 of everything absorbed so
//                  far (the state itself continues); the scalar is the 64
//                  digest bytes read as a little-endian integer mod r
// Blake2b is RFC 7693, written here from the RFC.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "guard.h"
#include "host_fr.h"
#include "zkwasm_prover_kernel.h"

namespace {

constexpr uint64_t kIV[8] = {0x6a09e667f3bcc908ull, 0xbb67ae8584caa73bull, 0x3c6ef372fe94f82bull,
                             0xa54ff53a5f1d36f1ull, 0x510e527fade682d1ull, 0x9b05688c2b3e6c1full,
                             0x1f83d9abfb41bd6bull, 0x5be0cd19137e2179ull};
constexpr uint8_t kSigma[12][16] = {
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}, {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3},
    {11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4}, {7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8},
    {9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13}, {2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9},
    {12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11}, {13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10},
    {6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5}, {10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0},
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}, {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3}};

inline uint64_t rotr(uint64_t x, int n) { return (x >> n) | (x << (64 - n)); }
inline uint64_t load64(const uint8_t *p) {
    uint64_t v = 0;
    for (int i = 7; i >= 0; i--) v = (v << 8) | p[i];
    return v;
}

struct Blake2b {
    uint64_t h[8];
    uint64_t t[2];
    uint8_t buf[128];
    size_t len;  // bytes in buf
    size_t outlen;

    void init(size_t out, const uint8_t *personal16) {
        outlen = out;
        for (int i = 0; i < 8; i++) h[i] = kIV[i];
        // parameter block: digest length, key length 0, fanout 1, depth 1;
        // personalisation at bytes 48..63
        h[0] ^= 0x01010000ull ^ (uint64_t)out;
        if (personal16) {
            h[6] ^= load64(personal16);
            h[7] ^= load64(personal16 + 8);
        }
        t[0] = t[1] = 0;
        len = 0;
    }
    void compress(const uint8_t *block, bool last) {
        uint64_t v[16], m[16];
        for (int i = 0; i < 8; i++) {
            v[i] = h[i];
            v[i + 8] = kIV[i];
        }
        v[12] ^= t[0];
        v[13] ^= t[1];
        if (last) v[14] = ~v[14];
        for (int i = 0; i < 16; i++) m[i] = load64(block + 8 * i);
        auto G = [&](int a, int b, int c, int d, uint64_t x, uint64_t y) {
            v[a] = v[a] + v[b] + x;
            v[d] = rotr(v[d] ^ v[a], 32);
            v[c] = v[c] + v[d];
            v[b] = rotr(v[b] ^ v[c], 24);
            v[a] = v[a] + v[b] + y;
            v[d] = rotr(v[d] ^ v[a], 16);
            v[c] = v[c] + v[d];
            v[b] = rotr(v[b] ^ v[c], 63);
        };
        for (int r = 0; r < 12; r++) {
            const uint8_t *s = kSigma[r];
            G(0, 4, 8, 12, m[s[0]], m[s[1]]);
            G(1, 5, 9, 13, m[s[2]], m[s[3]]);
            G(2, 6, 10, 14, m[s[4]], m[s[5]]);
            G(3, 7, 11, 15, m[s[6]], m[s[7]]);
            G(0, 5, 10, 15, m[s[8]], m[s[9]]);
            G(1, 6, 11, 12, m[s[10]], m[s[11]]);
            G(2, 7, 8, 13, m[s[12]], m[s[13]]);
            G(3, 4, 9, 14, m[s[14]], m[s[15]]);
        }
        for (int i = 0; i < 8; i++) h[i] ^= v[i] ^ v[i + 8];
    }
    void update(const uint8_t *p, size_t n) {
        while (n) {
            if (len == 128) {  // a full block is compressed only once more input follows
                t[0] += 128;
                if (t[0] < 128) t[1]++;
                compress(buf, false);
                len = 0;
            }
            size_t k = 128 - len < n ? 128 - len : n;
            memcpy(buf + len, p, k);
            len += k;
            p += k;
            n -= k;
        }
    }
    void final(uint8_t *out) const {
        Blake2b c = *this;
        c.t[0] += c.len;
        if (c.t[0] < c.len) c.t[1]++;
        memset(c.buf + c.len, 0, 128 - c.len);
        c.compress(c.buf, true);
        for (size_t i = 0; i < outlen; i++) out[i] = (uint8_t)(c.h[i / 8] >> (8 * (i % 8)));
    }
};

const uint8_t kPersonal[16] = {'H', 'a', 'l', 'o', '2', '-', 'T', 'r', 'a', 'n', 's', 'c', 'r', 'i', 'p', 't'};

// 2^768 mod r in Montgomery form's terms: from_bytes_wide = lo * R^2 + hi * R^3
// under Montgomery products (lo, hi the 256-bit halves, read unreduced)
constexpr uint64_t kR3[4] = {0x5e94d8e1b4bf0040ull, 0x2a489cbe1cfbb6b8ull, 0x893cc664a19fcfedull,
                             0x0cf8594b7fcc657cull};

}  // namespace

extern "C" {

// Blake2b of in (RFC 7693), outlen in [1, 64], optional 16-byte personalisation.
int zk_blake2b(void *out, int outlen, const void *in, size_t inlen, const void *personal16) {
    if (!out || outlen < 1 || outlen > 64 || (inlen && !in)) return -1;
    Blake2b b;
    b.init((size_t)outlen, (const uint8_t *)personal16);
    b.update((const uint8_t *)in, inlen);
    b.final((uint8_t *)out);
    return 0;
}

void *zk_transcript_new(void) {
    return zk::guard_or<void *>(nullptr, [&]() -> void * {
        Blake2b *b = new Blake2b();
        b->init(64, kPersonal);
        return b;
    });
}
void zk_transcript_free(void *t) { delete (Blake2b *)t; }

// x, y: canonical little-endian coordinates (32 bytes each)
int zk_transcript_common_point(void *t, const void *x, const void *y) {
    if (!t || !x || !y) return -1;
    const uint8_t tag = 1;
    Blake2b *b = (Blake2b *)t;
    b->update(&tag, 1);
    b->update((const uint8_t *)x, 32);
    b->update((const uint8_t *)y, 32);
    return 0;
}
// s: canonical little-endian scalar (32 bytes)
int zk_transcript_common_scalar(void *t, const void *s) {
    if (!t || !s) return -1;
    const uint8_t tag = 2;
    Blake2b *b = (Blake2b *)t;
    b->update(&tag, 1);
    b->update((const uint8_t *)s, 32);
    return 0;
}
// 64 little-endian bytes reduced mod r (halo2's Fr::from_bytes_wide), as a
// Montgomery-form Fr (4 x u64) in mont_out
int zk_fr_from_bytes_wide(const void *in64, void *mont_out) {
    if (!in64 || !mont_out) return -1;
    const uint8_t *d = (const uint8_t *)in64;
    {
        zkh::Fr lo, hi;
        for (int i = 0; i < 4; i++) {
            uint64_t a = 0, c = 0;
            for (int k = 7; k >= 0; k--) {
                a = (a << 8) | d[8 * i + k];
                c = (c << 8) | d[32 + 8 * i + k];
            }
            lo.l[i] = a;
            hi.l[i] = c;
        }
        const zkh::Fr r2{{zkh::kR2[0], zkh::kR2[1], zkh::kR2[2], zkh::kR2[3]}};
        const zkh::Fr r3{{kR3[0], kR3[1], kR3[2], kR3[3]}};
        zkh::Fr v = zkh::fr_add(zkh::fr_mul(lo, r2), zkh::fr_mul(hi, r3));
        memcpy(mont_out, v.l, 32);
    }
    return 0;
}

// challenge: Montgomery-form Fr (4 x u64) in mont_out, the 64 digest bytes
// in digest64 (either may be null)
int zk_transcript_squeeze(void *t, void *mont_out, void *digest64) {
    if (!t) return -1;
    const uint8_t tag = 0;
    Blake2b *b = (Blake2b *)t;
    b->update(&tag, 1);
    uint8_t d[64];
    b->final(d);
    if (digest64) memcpy(digest64, d, 64);
    if (mont_out) zk_fr_from_bytes_wide(d, mont_out);
    return 0;
}

}  // extern "C"
