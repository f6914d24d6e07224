// Device kernels that synthesise the witness of the PLONKish benchmark
// circuit (gate a*b + c - d = 0, range lookups, copy cycles, shuffles) and the
// small helpers the prover core needs between library calls.  Rows at and
// past u = n - (blinding + 1) are blinding rows: random, unconstrained.
#include "zkcore.h"
#include "zkinternal.h"

using namespace zk;

namespace {

constexpr uint32_t kT = 256;

uint32_t blocks_for(uint64_t n) {
    uint64_t b = (n + kT - 1) / kT;
    return (uint32_t)(b < 65536 ? (b ? b : 1) : 65536);
}

#define ROWS(i, n)                                                                      \
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (uint32_t)(n); \
         i += gridDim.x * blockDim.x)

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ void k_u32_rows_to_fr(Fr *col, const uint32_t *__restrict__ v, uint32_t rows) {
    ROWS(i, rows) fe_store<FrParams>(&col[i], fe_from_u32<FrParams>(v[i]));
}

// lookup input values below u: hash(seed, i) mod tsize
__global__ void k_lookup_values(uint32_t *vals, uint32_t u, uint32_t tsize, uint64_t seed) {
    ROWS(i, u) vals[i] = (uint32_t)(mix64(seed ^ ((uint64_t)i * 0x9e3779b97f4a7c15ull)) % tsize);
}

// col[i] = i mod tsize for i < u (Montgomery)
__global__ void k_table_rows(Fr *col, uint32_t u, uint32_t tsize) {
    ROWS(i, u) fe_store<FrParams>(&col[i], fe_from_u32<FrParams>(i % tsize));
}

__global__ void k_gate_rows(Fr *d, const Fr *__restrict__ a, const Fr *__restrict__ b, const Fr *__restrict__ c,
                            uint32_t u) {
    ROWS(i, u) fe_store<FrParams>(&d[i], fe_load<FrParams>(&a[i]) * fe_load<FrParams>(&b[i]) + fe_load<FrParams>(&c[i]));
}

// d[i] = x[i] * y[i] for i < u (the fixed-coefficient gates' outputs)
__global__ void k_mul_rows(Fr *d, const Fr *__restrict__ x, const Fr *__restrict__ y, uint32_t u) {
    ROWS(i, u) fe_store<FrParams>(&d[i], fe_load<FrParams>(&x[i]) * fe_load<FrParams>(&y[i]));
}

// out[i] = 1 on [lo, hi), 0 elsewhere
__global__ void k_indicator(Fr *out, uint32_t n, uint32_t lo, uint32_t hi) {
    ROWS(i, n) fe_store<FrParams>(&out[i], (i >= lo && i < hi) ? fe_one<FrParams>() : fe_zero<FrParams>());
}

__global__ void k_zero_rows(Fr *col, uint32_t lo, uint32_t hi) {
    ROWS(k, hi - lo) fe_store<FrParams>(&col[lo + k], fe_zero<FrParams>());
}

__global__ void k_gather(Fr *dst, const Fr *__restrict__ src, const uint32_t *__restrict__ idx, uint32_t rows) {
    ROWS(i, rows) fe_store<FrParams>(&dst[i], fe_load<FrParams>(&src[idx[i]]));
}

// cells are numbered col * u + row over the usable rows; head[cell] is the
// first cell of its copy cycle (itself when the cell is in none)
__global__ void k_cycle_values(Fr *const *__restrict__ cols, const uint32_t *__restrict__ head, uint32_t ncols,
                               uint32_t u) {
    ROWS(cell, ncols * u) {
        uint32_t h = head[cell];
        if (h != cell) {
            uint32_t c = cell / u, i = cell % u, hc = h / u, hi = h % u;
            fe_store<FrParams>(&cols[c][i], fe_load<FrParams>(&cols[hc][hi]));
        }
    }
}

// sigma_c[i] = delta^col(t) * w^row(t), t = target[c * u + i] for i < u and
// t = (c, i) for the blinding rows
__global__ void k_sigma_rows(Fr *const *__restrict__ sig, const uint32_t *__restrict__ target,
                             const Fr *__restrict__ deltas, const Fr *__restrict__ omegas, uint32_t ncols, uint32_t n,
                             uint32_t u) {
    ROWS(cell, (uint64_t)ncols * n) {
        uint32_t c = cell / n, i = cell % n;
        uint32_t tc = c, ti = i;
        if (i < u) {
            uint32_t t = target[c * u + i];
            tc = t / u;
            ti = t % u;
        }
        fe_store<FrParams>(&sig[c][i], fe_load<FrParams>(&deltas[tc]) * fe_load<FrParams>(&omegas[ti]));
    }
}

// col[i] = v[0]
__global__ void k_fill_value(Fr *col, const Fr *__restrict__ v, uint32_t n) {
    ROWS(i, n) fe_store<FrParams>(&col[i], fe_load<FrParams>(v));
}

// flag[0] = 1 if any of col[lo..hi) is nonzero (flag must start at 0)
__global__ void k_any_nonzero(const Fr *__restrict__ col, uint64_t lo, uint64_t hi, uint32_t *flag) {
    for (uint64_t i = lo + blockIdx.x * blockDim.x + threadIdx.x; i < hi; i += (uint64_t)gridDim.x * blockDim.x)
        if (!fe_is_zero<FrParams>(fe_load<FrParams>(&col[i]))) atomicOr(flag, 1u);
}

}  // namespace

namespace zk {

hipError_t wit_u32_rows_to_fr(void *col, const uint32_t *v, uint32_t rows, hipStream_t s) {
    if (!rows) return hipSuccess;
    k_u32_rows_to_fr<<<blocks_for(rows), kT, 0, s>>>((Fr *)col, v, rows);
    return hipGetLastError();
}
hipError_t wit_lookup_values(uint32_t *vals, uint32_t u, uint32_t tsize, uint64_t seed, hipStream_t s) {
    if (!u || !tsize) return hipErrorInvalidValue;
    k_lookup_values<<<blocks_for(u), kT, 0, s>>>(vals, u, tsize, seed);
    return hipGetLastError();
}
hipError_t wit_table_rows(void *col, uint32_t u, uint32_t tsize, hipStream_t s) {
    if (!u || !tsize) return hipErrorInvalidValue;
    k_table_rows<<<blocks_for(u), kT, 0, s>>>((Fr *)col, u, tsize);
    return hipGetLastError();
}
hipError_t wit_gate_rows(void *d, const void *a, const void *b, const void *c, uint32_t u, hipStream_t s) {
    if (!u) return hipSuccess;
    k_gate_rows<<<blocks_for(u), kT, 0, s>>>((Fr *)d, (const Fr *)a, (const Fr *)b, (const Fr *)c, u);
    return hipGetLastError();
}
hipError_t wit_mul_rows(void *d, const void *x, const void *y, uint32_t u, hipStream_t s) {
    if (!u) return hipSuccess;
    k_mul_rows<<<blocks_for(u), kT, 0, s>>>((Fr *)d, (const Fr *)x, (const Fr *)y, u);
    return hipGetLastError();
}
hipError_t wit_indicator(void *out, uint32_t n, uint32_t lo, uint32_t hi, hipStream_t s) {
    k_indicator<<<blocks_for(n), kT, 0, s>>>((Fr *)out, n, lo, hi);
    return hipGetLastError();
}
hipError_t wit_zero_rows(void *col, uint32_t lo, uint32_t hi, hipStream_t s) {
    if (hi <= lo) return hipSuccess;
    k_zero_rows<<<blocks_for(hi - lo), kT, 0, s>>>((Fr *)col, lo, hi);
    return hipGetLastError();
}
hipError_t wit_gather(void *dst, const void *src, const uint32_t *idx, uint32_t rows, hipStream_t s) {
    if (!rows) return hipSuccess;
    k_gather<<<blocks_for(rows), kT, 0, s>>>((Fr *)dst, (const Fr *)src, idx, rows);
    return hipGetLastError();
}
hipError_t wit_cycle_values(void *const *cols_dev, const uint32_t *head, uint32_t ncols, uint32_t u, hipStream_t s) {
    k_cycle_values<<<blocks_for((uint64_t)ncols * u), kT, 0, s>>>((Fr *const *)cols_dev, head, ncols, u);
    return hipGetLastError();
}
hipError_t wit_sigma_rows(void *const *sig_dev, const uint32_t *target, const void *deltas, const void *omegas,
                          uint32_t ncols, uint32_t n, uint32_t u, hipStream_t s) {
    k_sigma_rows<<<blocks_for((uint64_t)ncols * n), kT, 0, s>>>((Fr *const *)sig_dev, target, (const Fr *)deltas,
                                                                 (const Fr *)omegas, ncols, n, u);
    return hipGetLastError();
}
hipError_t wit_fill_value(void *col, const void *v, uint32_t n, hipStream_t s) {
    k_fill_value<<<blocks_for(n), kT, 0, s>>>((Fr *)col, (const Fr *)v, n);
    return hipGetLastError();
}
hipError_t wit_any_nonzero(const void *col, uint64_t lo, uint64_t hi, uint32_t *flag, hipStream_t s) {
    if (hi <= lo) return hipSuccess;
    k_any_nonzero<<<blocks_for(hi - lo), kT, 0, s>>>((const Fr *)col, lo, hi, flag);
    return hipGetLastError();
}

}  // namespace zk
