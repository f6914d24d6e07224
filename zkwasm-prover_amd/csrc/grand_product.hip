// Permutation and shuffle grand products on the device.
//
// The reference computes these on host cores (src/permutation.rs:16-150,
// src/shuffle.rs:16-215) while the GPU runs the logup work.  Here each is
// three device steps:
//   1. den[i]  = prod_j (denominator terms)            (one pass over columns)
//   2. den     = 1/den                                  (batch inversion)
//   3. ratio[i]= den[i] * prod_j (numerator terms)
// followed by an exclusive prefix-product scan, z[0] = 1.  The reference
// splits the product into chunks with carried tails; field multiplication is
// exact and associative, so every z[i] equals the sequential product bit for
// bit.  Rows after the last usable row are filled with blinding values by the
// caller, as fill_random does in the reference.
#include "zkcore.h"

using namespace zk;

namespace zk {
size_t binv_scratch_elems(uint32_t n);
hipError_t batch_invert_ex(const Fr *x, const Fr *beta, Fr *out, int mode, uint32_t n, Fr *temp,
                           hipStream_t stream);
size_t fscan_scratch_elems(uint32_t n);
hipError_t field_scan(Fr *z, uint32_t n, bool mul, const Fr *carry, Fr *last_z, int64_t last_index, Fr *tot,
                      hipStream_t stream);
}  // namespace zk

namespace {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kRows = 8;  // rows per thread (interleaved by T)

// den[i] = prod_j (v_j[i] + beta * sigma_j[i] + gamma)   (src/permutation.rs:53-70)
__global__ void k_perm_den(Fr *__restrict__ den, const Fr *const *__restrict__ vals, const Fr *const *__restrict__ sig,
                           int ncols, const Fr *__restrict__ beta, const Fr *__restrict__ gamma, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t stride = gridDim.x * blockDim.x;
    Fr b = fe_load<FrParams>(beta), g = fe_load<FrParams>(gamma);
    for (; i < n; i += stride) {
        Fr acc = fe_one<FrParams>();
        for (int j = 0; j < ncols; j++) {
            Fr t = b * fe_load<FrParams>(&sig[j][i]) + g + fe_load<FrParams>(&vals[j][i]);
            acc = j ? acc * t : t;
        }
        fe_store<FrParams>(&den[i], acc);
    }
}

// den[i] *= prod_j (v_j[i] + beta * delta_start * delta^j * omega^i + gamma)
// (src/permutation.rs:79-95).  Thread t owns rows t + k*T; omega^t comes from
// square-and-multiply, then steps by omega^T.
__global__ void k_perm_num(Fr *__restrict__ den, const Fr *const *__restrict__ vals, int ncols,
                           const Fr *__restrict__ beta, const Fr *__restrict__ gamma,
                           const Fr *__restrict__ delta_start, const Fr *__restrict__ delta,
                           const Fr *__restrict__ omega, uint32_t n, uint32_t T) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    Fr w = fe_load<FrParams>(omega);
    Fr wt = fe_one<FrParams>(), base = w;
    for (uint32_t e = t; e; e >>= 1) {
        if (e & 1) wt = wt * base;
        base = fe_sqr<FrParams>(base);
    }
    Fr wT = fe_one<FrParams>();
    base = w;
    for (uint32_t e = T; e; e >>= 1) {
        if (e & 1) wT = wT * base;
        base = fe_sqr<FrParams>(base);
    }
    Fr g = fe_load<FrParams>(gamma);
    Fr bd0 = fe_load<FrParams>(beta) * fe_load<FrParams>(delta_start);
    Fr d = fe_load<FrParams>(delta);
    for (uint32_t i = t; i < n; i += T) {
        Fr acc = fe_load<FrParams>(&den[i]);
        Fr x = bd0 * wt;  // beta * delta_start * delta^j * omega^i, j = 0
        for (int j = 0; j < ncols; j++) {
            acc = acc * (x + g + fe_load<FrParams>(&vals[j][i]));
            if (j + 1 < ncols) x = x * d;
        }
        fe_store<FrParams>(&den[i], acc);
        wt = wt * wT;
    }
}

// den[i] = prod_j (shuffle_j[i] + betas[j])   (src/shuffle.rs:148-161)
__global__ void k_shuffle_den(Fr *__restrict__ den, const Fr *const *__restrict__ cols, const Fr *__restrict__ betas,
                              int group_len, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t stride = gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        Fr acc = fe_load<FrParams>(&cols[0][i]) + fe_load<FrParams>(&betas[0]);
        for (int j = 1; j < group_len; j++) acc = acc * (fe_load<FrParams>(&cols[j][i]) + fe_load<FrParams>(&betas[j]));
        fe_store<FrParams>(&den[i], acc);
    }
}
// den[i] *= prod_j (input_j[i] + betas[j])   (src/shuffle.rs:167-176)
__global__ void k_shuffle_num(Fr *__restrict__ den, const Fr *const *__restrict__ cols, const Fr *__restrict__ betas,
                              int group_len, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t stride = gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        Fr acc = fe_load<FrParams>(&den[i]);
        for (int j = 0; j < group_len; j++) acc = acc * (fe_load<FrParams>(&cols[j][i]) + fe_load<FrParams>(&betas[j]));
        fe_store<FrParams>(&den[i], acc);
    }
}

// tails[s] = prod_{s' < s} z_{s'}[u]   (src/permutation.rs:131-145); one thread
__global__ void k_perm_tails(const Fr *const *__restrict__ zs, int nsets, uint32_t u, Fr *__restrict__ tails) {
    if (threadIdx.x || blockIdx.x) return;
    Fr acc = fe_one<FrParams>();
    for (int s = 0; s < nsets; s++) {
        fe_store<FrParams>(&tails[s], acc);
        acc = acc * fe_load<FrParams>(&zs[s][u]);
    }
}
// z_s[i] *= tails[s] for i <= u, s >= 1
__global__ void k_perm_chain(Fr *const *__restrict__ zs, int nsets, uint32_t u, const Fr *__restrict__ tails) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t s = blockIdx.y + 1;
    if (i > u || s >= (uint32_t)nsets) return;
    fe_store<FrParams>(&zs[s][i], fe_load<FrParams>(&zs[s][i]) * fe_load<FrParams>(&tails[s]));
}

uint32_t grid_for(uint32_t n) {
    uint32_t b = (n + kThreads - 1) / kThreads;
    return b > 16384 ? 16384 : (b ? b : 1);
}

}  // namespace

extern "C" {

hipError_t zk_permutation_z_set(void *z, const void *values, const void *sigmas, int ncols, const void *beta,
                                const void *gamma, const void *delta_start, const void *delta, const void *omega,
                                void *temp, int n, hipStream_t stream) {
    if (n <= 0 || ncols <= 0 || !z || !values || !sigmas || !beta || !gamma || !delta_start || !delta || !omega ||
        !temp)
        return hipErrorInvalidValue;
    Fr *zz = (Fr *)z;
    Fr *tmp = (Fr *)temp;
    const uint32_t un = (uint32_t)n;
    k_perm_den<<<grid_for(un), kThreads, 0, stream>>>(zz, (const Fr *const *)values, (const Fr *const *)sigmas, ncols,
                                                      (const Fr *)beta, (const Fr *)gamma, un);
    ZK_CHECK(hipGetLastError());
    ZK_CHECK(batch_invert_ex(zz, nullptr, zz, 0, un, tmp, stream));
    uint32_t T = (un + kRows - 1) / kRows;
    k_perm_num<<<(T + kThreads - 1) / kThreads, kThreads, 0, stream>>>(
        zz, (const Fr *const *)values, ncols, (const Fr *)beta, (const Fr *)gamma, (const Fr *)delta_start,
        (const Fr *)delta, (const Fr *)omega, un, T);
    ZK_CHECK(hipGetLastError());
    return field_scan(zz, un, true, nullptr, nullptr, -1, tmp, stream);
}

hipError_t zk_permutation_chain(void *zs, int nsets, int u, int n, void *tails, hipStream_t stream) {
    if (nsets <= 0 || n <= 0 || u < 0 || u >= n || !zs || !tails) return hipErrorInvalidValue;
    k_perm_tails<<<1, 64, 0, stream>>>((const Fr *const *)zs, nsets, (uint32_t)u, (Fr *)tails);
    ZK_CHECK(hipGetLastError());
    if (nsets > 1) {
        dim3 grid(((uint32_t)u + 1 + kThreads - 1) / kThreads, (uint32_t)nsets - 1);
        k_perm_chain<<<grid, kThreads, 0, stream>>>((Fr *const *)zs, nsets, (uint32_t)u, (const Fr *)tails);
        ZK_CHECK(hipGetLastError());
    }
    return hipSuccess;
}

hipError_t zk_shuffle_z(void *z, const void *inputs, const void *shuffles, int group_len, const void *betas, void *temp,
                        int n, hipStream_t stream) {
    if (n <= 0 || group_len <= 0 || !z || !inputs || !shuffles || !betas || !temp) return hipErrorInvalidValue;
    Fr *zz = (Fr *)z;
    const uint32_t un = (uint32_t)n;
    k_shuffle_den<<<grid_for(un), kThreads, 0, stream>>>(zz, (const Fr *const *)shuffles, (const Fr *)betas, group_len,
                                                         un);
    ZK_CHECK(hipGetLastError());
    ZK_CHECK(batch_invert_ex(zz, nullptr, zz, 0, un, (Fr *)temp, stream));
    k_shuffle_num<<<grid_for(un), kThreads, 0, stream>>>(zz, (const Fr *const *)inputs, (const Fr *)betas, group_len,
                                                         un);
    ZK_CHECK(hipGetLastError());
    return field_scan(zz, un, true, nullptr, nullptr, -1, (Fr *)temp, stream);
}

}  // extern "C"
