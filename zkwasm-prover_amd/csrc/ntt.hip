// Fr NTT / iNTT / coset preparation for gfx950.
//
// Drop-in for `ntt`, `expand_omega_buffer` and `extended_prepare`
// (cuda/bn254.cu:1232-1253, :1333-1428; caller src/cuda/ntt.rs:15-135).
//
// Algorithm: Stockham auto-sort radix-2^d passes, natural order in and out.
// Pass p (log_p bits already done, this pass d bits, t = n >> d):
//   u[i]  = x[idx + i*t] * w^((n >> log_p >> d) * k * i),  k = idx mod 2^log_p
//   U     = DFT_{2^d}(u)                         (radix-2 DIF in LDS)
//   y[(idx - k) * 2^d + k + j * 2^log_p] = U[j]
// A workgroup owns a 1024-element tile: G = 1024 >> d consecutive indices by
// 2^d rows.  Each row of G consecutive elements is one coalesced HBM segment
// on load; the store order is chosen so that every wave writes contiguous
// runs as well.  Pass twiddles are read from the natural-order w^i table that
// `expand_omega_buffer` builds (one gather per element, no extra multiply);
// the 2^(d-1) in-pass twiddles are staged once per workgroup in LDS.  The
// n^-1 scaling of the iNTT is fused into the last pass.
#include "zkcore.h"

using namespace zk;

namespace {

constexpr uint32_t kTile = 1024;
constexpr uint32_t kThreads = 256;
constexpr uint32_t kMaxDeg = 8;

__global__ __launch_bounds__(kThreads) void k_ntt_pass(const Fr *__restrict__ x, Fr *__restrict__ y,
                                                       const Fr *__restrict__ omegas,
                                                       const Fr *__restrict__ n_inv, uint32_t log_n,
                                                       uint32_t log_p, uint32_t deg, uint32_t log_g) {
    extern __shared__ Fr lds[];
    const uint32_t tid = threadIdx.x, bs = blockDim.x;
    const uint32_t D = 1u << deg;
    const uint32_t G = 1u << log_g;
    const uint32_t tile = D << log_g;
    const uint32_t log_t = log_n - deg;
    const uint32_t p = 1u << log_p;
    const uint32_t index0 = blockIdx.x << log_g;
    Fr *s = lds;
    Fr *tw = lds + tile;

    for (uint32_t j = tid; j < D / 2; j += bs) tw[j] = fe_load<FrParams>(&omegas[j << log_t]);

    const uint32_t base_shift = log_t - log_p;
    for (uint32_t e = tid; e < tile; e += bs) {
        uint32_t g = e & (G - 1), i = e >> log_g;
        uint32_t index = index0 + g;
        Fr v = fe_load<FrParams>(&x[(size_t)index + ((size_t)i << log_t)]);
        if (log_p) {
            uint32_t k = index & (p - 1);
            uint32_t ex = (k * i) << base_shift;
            if (ex) v = v * fe_load<FrParams>(&omegas[ex]);
        }
        s[(g << deg) + i] = v;
    }
    __syncthreads();

    for (uint32_t rnd = 0; rnd < deg; rnd++) {
        const uint32_t bit = (D >> 1) >> rnd;
        for (uint32_t b = tid; b < tile / 2; b += bs) {
            uint32_t g = b >> (deg - 1);
            uint32_t ii = b & (D / 2 - 1);
            uint32_t di = ii & (bit - 1);
            uint32_t i0 = (ii << 1) - di;
            uint32_t i1 = i0 + bit;
            Fr a = s[(g << deg) + i0], c = s[(g << deg) + i1];
            s[(g << deg) + i0] = a + c;
            Fr d = a - c;
            if (di) d = d * tw[di << rnd];
            s[(g << deg) + i1] = d;
        }
        __syncthreads();
    }

    const uint32_t lp = log_p < log_g ? log_p : log_g;
    const bool scale = n_inv != nullptr;
    Fr ninv = scale ? fe_load<FrParams>(n_inv) : fe_zero<FrParams>();
    for (uint32_t e = tid; e < tile; e += bs) {
        uint32_t klow = e & ((1u << lp) - 1);
        uint32_t j = (e >> lp) & (D - 1);
        uint32_t hi = e >> (lp + deg);
        uint32_t g = (hi << lp) + klow;
        uint32_t index = index0 + g;
        uint32_t k = index & (p - 1);
        size_t addr = ((size_t)(index - k) << deg) + k + ((size_t)j << log_p);
        Fr v = s[(g << deg) + bitrev32(j, deg)];
        if (scale) v = v * ninv;
        fe_store<FrParams>(&y[addr], v);
    }
}

// buf[i] = w^i for i in [2, n) given buf[1] = w (cuda/bn254.cu:1073-1094)
__global__ void k_expand_omega(Fr *buf, uint32_t n, uint32_t chunk) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t start = (uint64_t)t * chunk;
    if (start >= n) return;
    uint64_t end = start + chunk < n ? start + chunk : n;
    Fr w = fe_load<FrParams>(&buf[1]);
    // w^start by square-and-multiply
    Fr acc = fe_one<FrParams>(), base = w;
    for (uint64_t e = start; e; e >>= 1) {
        if (e & 1) acc = acc * base;
        base = fe_sqr<FrParams>(base);
    }
    for (uint64_t i = start; i < end; i++) {
        fe_store<FrParams>(&buf[i], acc);
        acc = acc * w;
    }
}

__global__ void k_fill_zero(uint4 *p, size_t n16) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) p[i] = make_uint4(0, 0, 0, 0);
}

// s[i] *= coset_powers[(i % cn) - 1] for i % cn != 0 (cuda/bn254.cu:641-662)
__global__ void k_coset_scale(Fr *s, const Fr *__restrict__ cp, uint32_t cn, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t stride = gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        uint32_t r = i % cn;
        if (r) fe_store<FrParams>(&s[i], fe_load<FrParams>(&s[i]) * fe_load<FrParams>(&cp[r - 1]));
    }
}

}  // namespace

namespace zk {

// Radix plan: ceil(log_n / 8) passes, bits spread evenly (larger first).
int ntt_plan(uint32_t log_n, uint32_t *degs) {
    if (log_n == 0) return 0;
    uint32_t passes = (log_n + kMaxDeg - 1) / kMaxDeg;
    uint32_t rem = log_n;
    for (uint32_t i = 0; i < passes; i++) {
        uint32_t left = passes - i;
        degs[i] = (rem + left - 1) / left;
        rem -= degs[i];
    }
    return (int)passes;
}

// Runs the passes, ping-ponging between a and b; returns the number of
// passes (the result is in b when odd).
hipError_t ntt_run(Fr *a, Fr *b, const Fr *omegas, const Fr *n_inv, uint32_t log_n,
                   hipStream_t stream, int *passes_out) {
    uint32_t degs[8];
    int passes = ntt_plan(log_n, degs);
    Fr *src = a, *dst = b;
    uint32_t log_p = 0;
    const uint32_t n = 1u << log_n;
    for (int ps = 0; ps < passes; ps++) {
        uint32_t deg = degs[ps];
        uint32_t tile = n < kTile ? n : kTile;
        uint32_t log_tile = log2_exact(tile);
        uint32_t log_g = log_tile > deg ? log_tile - deg : 0;
        uint32_t tile_elems = (1u << deg) << log_g;
        uint32_t blocks = n / tile_elems;
        size_t lds = ((size_t)tile_elems + (1u << deg) / 2) * sizeof(Fr);
        k_ntt_pass<<<blocks, kThreads, lds, stream>>>(src, dst, omegas,
                                                      ps == passes - 1 ? n_inv : nullptr, log_n,
                                                      log_p, deg, log_g);
        ZK_CHECK(hipGetLastError());
        log_p += deg;
        Fr *t = src;
        src = dst;
        dst = t;
    }
    if (passes_out) *passes_out = passes;
    return hipSuccess;
}

hipError_t fill_zero(void *p, size_t bytes, hipStream_t stream) {
    if (!bytes) return hipSuccess;
    size_t n16 = bytes / 16;
    size_t blocks = (n16 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    k_fill_zero<<<(uint32_t)blocks, 256, 0, stream>>>((uint4 *)p, n16);
    return hipGetLastError();
}

hipError_t coset_scale(Fr *s, const Fr *cp, uint32_t cn, uint32_t n, hipStream_t stream) {
    if (!n) return hipSuccess;
    uint32_t blocks = (n + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    k_coset_scale<<<blocks, 256, 0, stream>>>(s, cp, cn, n);
    return hipGetLastError();
}

}  // namespace zk

extern "C" {

// cuda/bn254.cu:1357-1428.  buf holds the input in natural order; the output
// (natural order) lands in buf or tmp and *swap tells which (true: tmp).
// pq and max_deg are accepted for ABI compatibility; this kernel takes its
// in-pass twiddles from the w^i table that expand_omega_buffer produced.
hipError_t ntt(void *buf, void *tmp, const void *pq, const void *omegas, const void *n_inv,
               int log_n, int max_deg, bool *swap, hipStream_t stream) {
    (void)pq;
    (void)max_deg;
    if (log_n < 0 || log_n > 30 || !buf || !omegas) return hipErrorInvalidValue;
    if (log_n > 0 && !tmp) return hipErrorInvalidValue;
    int passes = 0;
    ZK_CHECK(ntt_run((Fr *)buf, (Fr *)tmp, (const Fr *)omegas, (const Fr *)n_inv, (uint32_t)log_n,
                     stream, &passes));
    if (swap) *swap = (passes & 1) != 0;
    return hipSuccess;
}

// cuda/bn254.cu:1333-1355.  buf[0] = 1 and buf[1] = w are set by the caller
// (src/cuda/ntt.rs:21-23); fills buf[i] = w^i for all i < 2^log_n in natural
// order.  The reference stores a bit-reversed half table for log_n in {22, 24}
// because its merge kernel indexes it that way; `ntt` here reads the natural
// table for every size, so may_bit_reversed does not change the layout.
hipError_t expand_omega_buffer(void *buf, int log_n, int may_bit_reversed) {
    (void)may_bit_reversed;
    if (log_n < 0 || log_n > 30 || !buf) return hipErrorInvalidValue;
    uint32_t n = 1u << log_n;
    if (n <= 2) return hipSuccess;
    uint32_t chunk = n >= (1u << 16) ? 64 : 8;
    uint32_t threads = (n + chunk - 1) / chunk;
    k_expand_omega<<<(threads + 63) / 64, 64, 0, 0>>>((Fr *)buf, n, chunk);
    return hipGetLastError();
}

// cuda/bn254.cu:1232-1253
hipError_t extended_prepare(void *s, const void *coset_powers, unsigned coset_powers_n, int size,
                            int extended_size, int to_coset, hipStream_t stream) {
    if (!s || size < 0 || extended_size < size || (coset_powers_n > 1 && !coset_powers) ||
        coset_powers_n == 0)
        return hipErrorInvalidValue;
    Fr *f = (Fr *)s;
    if (to_coset) return coset_scale(f, (const Fr *)coset_powers, coset_powers_n, (uint32_t)extended_size, stream);
    ZK_CHECK(fill_zero(f + size, (size_t)(extended_size - size) * sizeof(Fr), stream));
    return coset_scale(f, (const Fr *)coset_powers, coset_powers_n, (uint32_t)size, stream);
}

}  // extern "C"
