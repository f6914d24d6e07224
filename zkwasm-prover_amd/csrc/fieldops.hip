// Element-wise field / curve entry points used by the parity tests, plus the
// calibration kernels whose timings give the VALU roofline (Montgomery
// multiplies per second, single-wave latency of a multiply, a doubling and a
// mixed add) and the HBM copy baseline used to scale rocprof byte counters.
#include "ff29.h"
#include "zkcore.h"

using namespace zk;

namespace {

template <class P>
__global__ void k_fe_binop(Fe<P> *out, const Fe<P> *a, const Fe<P> *b, int op, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Fe<P> x = fe_load<P>(&a[i]);
    Fe<P> y = b ? fe_load<P>(&b[i]) : fe_zero<P>();
    Fe<P> r;
    switch (op) {
    case 0: r = fe_add<P>(x, y); break;
    case 1: r = fe_mul<P>(x, y); break;
    case 2: r = fe_sub<P>(x, y); break;
    case 3: r = fe_inv<P>(x); break;
    case 4: r = fe_sqr<P>(x); break;
    case 5: r = fe_neg<P>(x); break;
    case 6: r = fe_to_mont<P>(x); break;
    default: r = fe_from_mont<P>(x); break;
    }
    fe_store<P>(&out[i], r);
}

// The 29-bit-limb field through its conversions: op 0 a*b, 1 a+b, 2 a-b,
// 3 a^2, 4 round trip of a; inputs and outputs in the 32-bit Montgomery form.
__global__ void k_fp29_op(Fp *out, const Fp *a, const Fp *b, int op, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const F29 one = f29_one();
    F29 x = f29_mul(f29_from32(fe_load<FpParams>(&a[i])), one);  // reduced, < 2p
    F29 y = b ? f29_mul(f29_from32(fe_load<FpParams>(&b[i])), one) : x;
    F29 r;
    switch (op) {
    case 0: r = f29_mul(x, y); break;
    case 1: r = f29_add(x, y); break;
    case 2: r = f29_sub(x, y); break;
    case 3: r = f29_sqr(x); break;
    default: r = x; break;
    }
    fe_store<FpParams>(&out[i], f29_to32(r));
}

// out[i] = a[i] + b[i] on G1, with a, b affine; op selects the formula path
// 0: XYZZ mixed add -> Jacobian, 1: Jacobian add, 2: XYZZ dbl of a,
// 3: Jacobian dbl of a, 4: normalise(a + b) (z = 1)
__global__ __launch_bounds__(64) void k_g1_op(G1Jac *out, const G1Aff *a, const G1Aff *b, int op, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    G1Aff pa = aff_load(&a[i]);
    G1Aff pb = aff_load(&b[i]);
    G1Xyzz xa = xyzz_from_aff(pa, false);
    G1Jac r;
    if (op == 0) {
        G1Xyzz s = aff_is_id(pb) ? xa : xyzz_add_aff(xa, pb.x, pb.y);
        r = xyzz_to_jac(s);
    } else if (op == 1) {
        G1Jac ja = xyzz_to_jac(xa), jb = xyzz_to_jac(xyzz_from_aff(pb, false));
        r = jac_add(ja, jb);
    } else if (op == 2) {
        r = xyzz_to_jac(xyzz_dbl(xa));
    } else if (op == 3) {
        r = jac_dbl(xyzz_to_jac(xa));
    } else {
        G1Xyzz s = xyzz_add(xa, xyzz_from_aff(pb, false));
        r = jac_normalize(xyzz_to_jac(s));
    }
    jac_store(&out[i], r);
}

// throughput: each thread runs `iters` dependent multiplies of its own element
__global__ void k_bench_mul(Fr *buf, uint32_t iters) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    Fr x = fe_load<FrParams>(&buf[i]);
    Fr y = fe_load<FrParams>(&buf[i ^ 1]);
    for (uint32_t k = 0; k < iters; k++) {
        x = fe_mul<FrParams>(x, y);
        y = fe_mul<FrParams>(y, x);
    }
    fe_store<FrParams>(&buf[i], x);
}
// 29-bit-limb multiply throughput: variant 0 one chain per column, 1 split chains
template <int V>
__global__ void k_bench_mul29(Fp *buf, uint32_t iters) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const F29 one = f29_one();
    F29 x = f29_mul(f29_from32(fe_load<FpParams>(&buf[i])), one);
    F29 y = f29_mul(f29_from32(fe_load<FpParams>(&buf[i ^ 1])), one);
    for (uint32_t k = 0; k < iters; k++) {
        if (V == 0) {
            x = f29_mul_chain(x, y);
            y = f29_mul_chain(y, x);
        } else {
            x = f29_mul(x, y);
            y = f29_mul(y, x);
        }
    }
    fe_store<FpParams>(&buf[i], f29_to32(x));
}
__global__ void k_bench_mul_cios(Fr *buf, uint32_t iters) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    Fr x = fe_load<FrParams>(&buf[i]);
    Fr y = fe_load<FrParams>(&buf[i ^ 1]);
    for (uint32_t k = 0; k < iters; k++) {
        x = fe_mul_cios<FrParams>(x, y);
        y = fe_mul_cios<FrParams>(y, x);
    }
    fe_store<FrParams>(&buf[i], x);
}
__global__ void k_bench_dbl(G1Jac *buf, uint32_t iters) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    G1Jac p = jac_load(&buf[i]);
    for (uint32_t k = 0; k < iters; k++) p = jac_dbl(p);
    jac_store(&buf[i], p);
}
__global__ void k_bench_madd(G1Xyzz *acc, const G1Aff *pts, uint32_t npts, uint32_t iters) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    G1Xyzz a = xyzz_load(&acc[i]);
    for (uint32_t k = 0; k < iters; k++) {
        G1Aff q = aff_load(&pts[(i * 7 + k * 131) % npts]);
        a = xyzz_add_aff(a, q.x, q.y);
    }
    xyzz_store(&acc[i], a);
}
__global__ void k_copy(uint4 *dst, const uint4 *src, size_t n16) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) dst[i] = src[i];
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// counter-based uniform value < r: element i draws 4 x u64 from a SplitMix64
// stream seeded by (seed, i), top limb masked to 254 bits, reduced once
__device__ __forceinline__ Fr rand_canonical(uint64_t seed, uint64_t i) {
    uint64_t st = seed ^ (i * 0xd1342543de82ef95ull + 0x632be59bd9b4e019ull);
    Fr c;
    for (int k = 0; k < 4; k++) {
        uint64_t v = splitmix64(st);
        c.v[2 * k] = (uint32_t)v;
        c.v[2 * k + 1] = (uint32_t)(v >> 32);
    }
    c.v[7] &= 0x3fffffffu;
    return fe_reduce_once<FrParams>(fe_reduce_once<FrParams>(c));
}

__global__ void k_rand_fr(Fr *out, uint32_t n, uint64_t seed) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fe_store<FrParams>(&out[i], fe_to_mont<FrParams>(rand_canonical(seed, i)));
}

// out[i] = k_i * G (affine, Montgomery) with k_i = rand_canonical(seed, i)
// k G (canonical k < 2^254, G = (1, 2)) as an affine point, by double-and-add
__device__ G1Aff g_times(const Fr &k) {
    Fp gx = fe_from_u32<FpParams>(1), gy = fe_from_u32<FpParams>(2);
    G1Xyzz acc = xyzz_id();
    for (int b = 253; b >= 0; b--) {
        acc = xyzz_dbl(acc);
        if ((k.v[b >> 5] >> (b & 31)) & 1u) acc = xyzz_add_aff(acc, gx, gy);
    }
    G1Jac j = jac_normalize(xyzz_to_jac(acc));
    G1Aff r;
    if (fe_is_zero<FpParams>(j.z)) {
        r.x = fe_zero<FpParams>();
        r.y = fe_zero<FpParams>();
    } else {
        r.x = j.x;
        r.y = j.y;
    }
    return r;
}

__global__ __launch_bounds__(64) void k_gen_points(G1Aff *out, uint32_t n, uint64_t seed) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const G1Aff r = g_times(rand_canonical(seed, i));
    fe_store<FpParams>(&out[i].x, r.x);
    fe_store<FpParams>(&out[i].y, r.y);
}

// out[i] = s[i] G for Montgomery scalars s
__global__ __launch_bounds__(64) void k_g_mul_scalars(G1Aff *out, const Fr *__restrict__ s, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const G1Aff r = g_times(fe_from_mont<FrParams>(fe_load<FrParams>(&s[i])));
    fe_store<FpParams>(&out[i].x, r.x);
    fe_store<FpParams>(&out[i].y, r.y);
}

}  // namespace

extern "C" {

// Seeded synthetic inputs for the bench (not the oracle's SplitMix64 stream).
hipError_t zk_rand_fr(void *out, int n, unsigned long long seed, hipStream_t stream) {
    if (n <= 0 || !out) return n == 0 ? hipSuccess : hipErrorInvalidValue;
    k_rand_fr<<<(n + 255) / 256, 256, 0, stream>>>((Fr *)out, (uint32_t)n, seed);
    return hipGetLastError();
}
hipError_t zk_gen_points(void *out, int n, unsigned long long seed, hipStream_t stream) {
    if (n <= 0 || !out) return n == 0 ? hipSuccess : hipErrorInvalidValue;
    k_gen_points<<<(n + 63) / 64, 64, 0, stream>>>((G1Aff *)out, (uint32_t)n, seed);
    return hipGetLastError();
}

hipError_t zk_g1_mul_scalars(void *out, const void *scalars, int n, hipStream_t stream) {
    if (n <= 0 || !out || !scalars) return n == 0 ? hipSuccess : hipErrorInvalidValue;
    k_g_mul_scalars<<<(n + 63) / 64, 64, 0, stream>>>((G1Aff *)out, (const Fr *)scalars, (uint32_t)n);
    return hipGetLastError();
}

hipError_t zk_fr_op(void *out, const void *a, const void *b, int op, int n, hipStream_t stream) {
    if (n <= 0) return n == 0 ? hipSuccess : hipErrorInvalidValue;
    if (!out || !a || op < 0 || op > 7) return hipErrorInvalidValue;
    k_fe_binop<FrParams><<<(n + 255) / 256, 256, 0, stream>>>((Fr *)out, (const Fr *)a, (const Fr *)b, op, n);
    return hipGetLastError();
}
hipError_t zk_fp_op(void *out, const void *a, const void *b, int op, int n, hipStream_t stream) {
    if (n <= 0) return n == 0 ? hipSuccess : hipErrorInvalidValue;
    if (!out || !a || op < 0 || op > 7) return hipErrorInvalidValue;
    k_fe_binop<FpParams><<<(n + 255) / 256, 256, 0, stream>>>((Fp *)out, (const Fp *)a, (const Fp *)b, op, n);
    return hipGetLastError();
}
hipError_t zk_fp29_op(void *out, const void *a, const void *b, int op, int n, hipStream_t stream) {
    if (n <= 0) return n == 0 ? hipSuccess : hipErrorInvalidValue;
    if (!out || !a || op < 0 || op > 4) return hipErrorInvalidValue;
    k_fp29_op<<<(n + 255) / 256, 256, 0, stream>>>((Fp *)out, (const Fp *)a, (const Fp *)b, op, (uint32_t)n);
    return hipGetLastError();
}
hipError_t zk_g1_op(void *out, const void *a, const void *b, int op, int n, hipStream_t stream) {
    if (n <= 0) return n == 0 ? hipSuccess : hipErrorInvalidValue;
    if (!out || !a || !b || op < 0 || op > 4) return hipErrorInvalidValue;  // every op reads b
    k_g1_op<<<(n + 63) / 64, 64, 0, stream>>>((G1Jac *)out, (const G1Aff *)a, (const G1Aff *)b, op, n);
    return hipGetLastError();
}
// a calibration launch: threads / block workgroups of block threads, each
// thread running iters dependent steps on its own element of buf
static bool bench_shape_ok(const void *buf, int threads, int block, int iters) {
    return buf && threads > 0 && block > 0 && block <= 1024 && threads % block == 0 && iters >= 0;
}
hipError_t zk_bench_mul(void *buf, int threads, int block, int iters, hipStream_t stream) {
    if (!bench_shape_ok(buf, threads, block, iters)) return hipErrorInvalidValue;
    k_bench_mul<<<threads / block, block, 0, stream>>>((Fr *)buf, iters);
    return hipGetLastError();
}
hipError_t zk_bench_mul29(void *buf, int threads, int block, int iters, int variant, hipStream_t stream) {
    if (!bench_shape_ok(buf, threads, block, iters)) return hipErrorInvalidValue;
    if (variant)
        k_bench_mul29<1><<<threads / block, block, 0, stream>>>((Fp *)buf, iters);
    else
        k_bench_mul29<0><<<threads / block, block, 0, stream>>>((Fp *)buf, iters);
    return hipGetLastError();
}
hipError_t zk_bench_mul_cios(void *buf, int threads, int block, int iters, hipStream_t stream) {
    if (!bench_shape_ok(buf, threads, block, iters)) return hipErrorInvalidValue;
    k_bench_mul_cios<<<threads / block, block, 0, stream>>>((Fr *)buf, iters);
    return hipGetLastError();
}
hipError_t zk_bench_dbl(void *buf, int threads, int block, int iters, hipStream_t stream) {
    if (!bench_shape_ok(buf, threads, block, iters)) return hipErrorInvalidValue;
    k_bench_dbl<<<threads / block, block, 0, stream>>>((G1Jac *)buf, iters);
    return hipGetLastError();
}
hipError_t zk_bench_madd(void *acc, const void *pts, int npts, int threads, int block, int iters,
                         hipStream_t stream) {
    if (!bench_shape_ok(acc, threads, block, iters) || !pts || npts <= 0) return hipErrorInvalidValue;
    k_bench_madd<<<threads / block, block, 0, stream>>>((G1Xyzz *)acc, (const G1Aff *)pts, npts, iters);
    return hipGetLastError();
}
hipError_t zk_copy(void *dst, const void *src, size_t bytes, hipStream_t stream) {
    if (bytes % 16 || (bytes && (!dst || !src))) return hipErrorInvalidValue;
    if (!bytes) return hipSuccess;
    size_t n16 = bytes / 16;
    k_copy<<<4096, 256, 0, stream>>>((uint4 *)dst, (const uint4 *)src, n16);
    return hipGetLastError();
}

}  // extern "C"
