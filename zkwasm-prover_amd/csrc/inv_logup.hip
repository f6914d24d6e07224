// Batch inversion, logup grand sums and field prefix scans for gfx950.
//
// Drop-in for `logup_sum_input_inv`, `eval_logup_z_compose` and
// `eval_logup_z_pure` (cuda/bn254.cu:19-220, :1914-1993; caller
// src/lib.rs:590-745), plus zk_batch_invert.
//
// Batch inversion is Montgomery's trick in three launches over tiles of 4096
// elements (prefix products into the caller's temp, every tile's total
// inverted by its own thread of a small launch, then two products per
// element on the way back; details at k_binv_up below).  Above 2^24 elements
// a multi-level hierarchy takes over: level 0 thread t owns the m elements
// t, t + T, t + 2T, ... (T = n/m, coalesced) and writes only their product;
// the T products are inverted recursively the same way, the last <= 512 by
// one workgroup with a product tree in LDS; on the way down each thread
// rebuilds its prefix products in registers and emits the inverses.
// A zero input stays zero and does not disturb the others (halo2
// BatchInvert semantics); the reference's chunked kernel would zero its chunk
// instead, a case its callers never produce (beta is random).
//
// The logup z scan (exclusive prefix sum with a carried start value) is a
// three-kernel block scan: block totals, one-block scan of the totals with
// the carry, then the in-block exclusive scan plus offset.
#include "zkcore.h"

using namespace zk;

namespace {

constexpr int kM = 8;          // elements per thread at each level
constexpr uint32_t kTop = 512;  // values inverted by the single-workgroup tree
constexpr uint32_t kScanBlock = 2048;  // elements per scan block (256 threads x 8)

enum ApplyMode { OUT_SET = 0, OUT_ADD = 1 };

__device__ __forceinline__ Fr load_plus(const Fr *x, size_t i, const Fr *beta, bool has_beta, const Fr &b) {
    Fr v = fe_load<FrParams>(&x[i]);
    return has_beta ? v + b : v;
}

// fe_inv with the binary GCD inlined, so that its limb arrays stay in
// registers (fe_inv's out-of-line call keeps them in scratch memory); the
// exact one-bit extended Euclid should the product check fail, as fe_inv.
// 0 -> 0.
__device__ __forceinline__ Fr fr_inv_inline(const Fr &a) {
    if (fe_is_zero<FrParams>(a)) return fe_zero<FrParams>();
    uint32_t v[8];
    const bool ok = zkinv::bingcd(v, a.v, FrParams::M, FrParams::INV);
    const Fr r = fe_from_raw<FrParams>(v) * fe_from_raw<FrParams>(FrParams::INVC);
    if (ok && fe_eq<FrParams>(a * r, fe_one<FrParams>())) return r;
    return fe_inv_raw<FrParams>(a) * fe_from_raw<FrParams>(FrParams::R3);
}

// tot[t] = prod_k x'[t + k*T] (zeros skipped)
__global__ __launch_bounds__(256) void k_binv_totals(const Fr *__restrict__ x, const Fr *__restrict__ beta, Fr *__restrict__ tot,
                              uint32_t n, uint32_t T) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const bool hb = beta != nullptr;
    Fr b = hb ? fe_load<FrParams>(beta) : fe_zero<FrParams>();
    Fr acc = fe_one<FrParams>();
    for (uint32_t i = t; i < n; i += T) {
        Fr v = load_plus(x, i, beta, hb, b);
        if (!fe_is_zero<FrParams>(v)) acc = acc * v;
    }
    fe_store<FrParams>(&tot[t], acc);
}

// inverse of each element from the inverse of its thread's product
template <int MODE>
__global__ __launch_bounds__(256) void k_binv_apply(const Fr *__restrict__ x, const Fr *__restrict__ beta,
                             const Fr *__restrict__ inv_tot, Fr *out, uint32_t n, uint32_t T) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const bool hb = beta != nullptr;
    Fr b = hb ? fe_load<FrParams>(beta) : fe_zero<FrParams>();
    Fr v[kM], pre[kM];
    Fr acc = fe_one<FrParams>();
#pragma unroll 8
    for (int k = 0; k < kM; k++) {
        uint32_t i = t + (uint32_t)k * T;
        v[k] = i < n ? load_plus(x, i, beta, hb, b) : fe_zero<FrParams>();
        pre[k] = acc;
        if (!fe_is_zero<FrParams>(v[k])) acc = acc * v[k];
    }
    Fr inv = fe_load<FrParams>(&inv_tot[t]);
#pragma unroll 8
    for (int k = kM - 1; k >= 0; k--) {
        uint32_t i = t + (uint32_t)k * T;
        Fr r;
        if (fe_is_zero<FrParams>(v[k])) {
            r = fe_zero<FrParams>();
        } else {
            r = inv * pre[k];
            inv = inv * v[k];
        }
        if (i < n) {
            if (MODE == OUT_ADD) r = fe_load<FrParams>(&out[i]) + r;
            fe_store<FrParams>(&out[i], r);
        }
    }
}

// one workgroup: inverses of x'[i] = x[i] (+ beta) for i < N <= kTop (zeros stay
// zero), stored (MODE OUT_SET) or added (OUT_ADD) to out
template <int MODE>
__global__ __launch_bounds__(256) void k_binv_top(const Fr *__restrict__ x, const Fr *__restrict__ beta, Fr *out,
                                                  uint32_t N) {
    __shared__ Fr tree[2 * kTop];
    __shared__ uint32_t zero[kTop];
    const uint32_t P = kTop;  // leaves at [P, 2P)
    const bool hb = beta != nullptr;
    Fr b = hb ? fe_load<FrParams>(beta) : fe_zero<FrParams>();
    for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
        Fr v = i < N ? load_plus(x, i, beta, hb, b) : fe_one<FrParams>();
        bool z = fe_is_zero<FrParams>(v);
        zero[i] = z;
        tree[P + i] = z ? fe_one<FrParams>() : v;
    }
    __syncthreads();
    for (uint32_t w = P / 2; w >= 1; w >>= 1) {
        for (uint32_t i = threadIdx.x; i < w; i += blockDim.x) tree[w + i] = tree[2 * (w + i)] * tree[2 * (w + i) + 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) tree[1] = fr_inv_inline(tree[1]);
    __syncthreads();
    for (uint32_t w = 1; w < P; w <<= 1) {
        for (uint32_t i = threadIdx.x; i < w; i += blockDim.x) {
            uint32_t p = w + i;
            Fr ip = tree[p];
            Fr l = tree[2 * p], r = tree[2 * p + 1];
            tree[2 * p] = ip * r;
            tree[2 * p + 1] = ip * l;
        }
        __syncthreads();
    }
    for (uint32_t i = threadIdx.x; i < N; i += blockDim.x) {
        Fr r = zero[i] ? fe_zero<FrParams>() : tree[P + i];
        if (MODE == OUT_ADD) r = fe_load<FrParams>(&out[i]) + r;
        fe_store<FrParams>(&out[i], r);
    }
}

// ---- batch inversion in three launches (n <= 2^24) ------------------------
// Tile b = 4096 elements, 256 threads x 16; thread t owns the elements
// b * 4096 + k * 256 + t (k < 16), so every load and store is coalesced.
//   up:   each thread's exclusive prefix products go to pre[] (the slot of its
//         first element is not needed: that prefix is 1), the 256 thread
//         products into an LDS product tree; the tile's total goes into the
//         free slot pre[b * 4096].
//   tiles: one thread per tile inverts its total in place (T independent
//         binary-GCD inversions side by side, no product tree over tiles).
//   down: each tile rebuilds its thread products (pre of the last element
//         times that element), the tree, walks it down from the tile's
//         inverse, and each thread emits its elements' inverses backwards
//         (two products per element).
// Three Montgomery products per element in all; traffic 5 x 32 B per element
// (x twice, pre written and read, out written).
constexpr uint32_t kBT = 256;
constexpr int kBM = 16;
constexpr uint32_t kTile = kBT * kBM;
constexpr uint32_t kTopMax = kBT * kBM;  // tiles up to which the tiled path runs (n <= 2^24)

// LDS product tree over kBT leaves at [kBT, 2 kBT): tree[1] = their product
__device__ __forceinline__ void tree_up(Fr *tree) {
    for (uint32_t w = kBT / 2; w >= 1; w >>= 1) {
        if (threadIdx.x < w) {
            const uint32_t p = w + threadIdx.x;
            tree[p] = tree[2 * p] * tree[2 * p + 1];
        }
        __syncthreads();
    }
}
// from tree[1] = the inverse of the product: each leaf becomes its inverse
__device__ __forceinline__ void tree_down(Fr *tree) {
    for (uint32_t w = 1; w < kBT; w <<= 1) {
        if (threadIdx.x < w) {
            const uint32_t p = w + threadIdx.x;
            const Fr ip = tree[p], l = tree[2 * p], r = tree[2 * p + 1];
            tree[2 * p] = ip * r;
            tree[2 * p + 1] = ip * l;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBT) void k_binv_up(const Fr *__restrict__ x, const Fr *__restrict__ beta,
                                                 Fr *__restrict__ pre, uint32_t n) {
    __shared__ Fr tree[2 * kBT];
    const uint32_t t = threadIdx.x, base = blockIdx.x * kTile;
    const bool hb = beta != nullptr;
    const Fr b = hb ? fe_load<FrParams>(beta) : fe_zero<FrParams>();
    Fr acc = fe_one<FrParams>();
#pragma unroll 4
    for (int k = 0; k < kBM; k++) {
        const uint32_t i = base + (uint32_t)k * kBT + t;
        if (i >= n) break;
        const Fr v = load_plus(x, i, beta, hb, b);
        if (k) fe_store<FrParams>(&pre[i], acc);
        if (!fe_is_zero<FrParams>(v)) acc = acc * v;
    }
    tree[kBT + t] = acc;
    __syncthreads();
    tree_up(tree);
    if (t == 0) fe_store<FrParams>(&pre[base], tree[1]);
}

// the T tile totals at pre[b * kTile] -> their inverses, in place: one
// independent inversion per thread (a separate launch keeps the inversion's
// registers out of the streaming kernels)
__global__ __launch_bounds__(64) void k_binv_tiles(Fr *pre, uint32_t T) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= T) return;
    Fr *p = &pre[(size_t)b * kTile];
    fe_store<FrParams>(p, fr_inv_inline(fe_load<FrParams>(p)));
}

template <int MODE>
__global__ __launch_bounds__(kBT) void k_binv_down(const Fr *x, const Fr *__restrict__ beta,
                                                   const Fr *__restrict__ pre, Fr *out, uint32_t n) {
    __shared__ Fr tree[2 * kBT];
    const uint32_t t = threadIdx.x, base = blockIdx.x * kTile;
    const bool hb = beta != nullptr;
    const Fr b = hb ? fe_load<FrParams>(beta) : fe_zero<FrParams>();
    // this thread's elements: k < kl
    int kl = 0;
    while (kl < kBM && base + (uint32_t)kl * kBT + t < n) kl++;
    Fr prod = fe_one<FrParams>();
    if (kl) {
        const uint32_t i = base + (uint32_t)(kl - 1) * kBT + t;
        const Fr v = load_plus(x, i, beta, hb, b);
        if (kl > 1) prod = fe_load<FrParams>(&pre[i]);
        if (!fe_is_zero<FrParams>(v)) prod = prod * v;
    }
    tree[kBT + t] = prod;
    __syncthreads();
    const Fr inv_tile = fe_load<FrParams>(&pre[base]);
    tree_up(tree);
    if (t == 0) tree[1] = inv_tile;
    __syncthreads();
    tree_down(tree);
    Fr acc = tree[kBT + t];
    for (int k = kl - 1; k >= 0; k--) {
        const uint32_t i = base + (uint32_t)k * kBT + t;
        const Fr v = load_plus(x, i, beta, hb, b);
        Fr r = fe_zero<FrParams>();
        if (!fe_is_zero<FrParams>(v)) {
            r = k ? acc * fe_load<FrParams>(&pre[i]) : acc;
            acc = acc * v;
        }
        if (MODE == OUT_ADD) r = fe_load<FrParams>(&out[i]) + r;
        fe_store<FrParams>(&out[i], r);
    }
}

__global__ void k_logup_step2(Fr *__restrict__ z, const Fr *__restrict__ s, const Fr *__restrict__ tinv,
                              const Fr *__restrict__ m, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t stride = gridDim.x * blockDim.x;
    for (; i < n; i += stride)
        fe_store<FrParams>(&z[i], fe_load<FrParams>(&s[i]) - fe_load<FrParams>(&m[i]) * fe_load<FrParams>(&tinv[i]));
}

// ---- field prefix scans (add or mul) --------------------------------------
template <bool MUL>
__device__ __forceinline__ Fr op_combine(const Fr &a, const Fr &b) {
    return MUL ? a * b : a + b;
}
template <bool MUL>
__device__ __forceinline__ Fr op_identity() {
    return MUL ? fe_one<FrParams>() : fe_zero<FrParams>();
}

// block totals: 256 threads x 8 contiguous elements
template <bool MUL>
__global__ __launch_bounds__(256) void k_fscan_totals(const Fr *__restrict__ z, uint32_t n, Fr *__restrict__ tot) {
    __shared__ Fr sh[256];
    uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * 8;
    Fr acc = op_identity<MUL>();
    for (uint32_t k = 0; k < 8; k++)
        if (base + k < n) acc = op_combine<MUL>(acc, fe_load<FrParams>(&z[base + k]));
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t s = 128; s; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = op_combine<MUL>(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) fe_store<FrParams>(&tot[blockIdx.x], sh[0]);
}

// one block: exclusive scan of nb totals in place, starting from *carry
// (identity when carry == null)
template <bool MUL>
__global__ __launch_bounds__(256) void k_fscan_top(Fr *tot, uint32_t nb, const Fr *__restrict__ carry) {
    __shared__ Fr sh[256];
    __shared__ Fr run;
    if (threadIdx.x == 0) run = carry ? fe_load<FrParams>(carry) : op_identity<MUL>();
    __syncthreads();
    // each round covers 256 * 8 totals: thread-local serial part + LDS scan
    for (uint32_t base = 0; base < nb; base += 256 * 8) {
        uint32_t b0 = base + threadIdx.x * 8;
        Fr loc[8];
        Fr acc = op_identity<MUL>();
#pragma unroll 8
        for (uint32_t k = 0; k < 8; k++) {
            loc[k] = b0 + k < nb ? fe_load<FrParams>(&tot[b0 + k]) : op_identity<MUL>();
            acc = op_combine<MUL>(acc, loc[k]);
        }
        sh[threadIdx.x] = acc;
        __syncthreads();
        // Hillis-Steele inclusive scan of the 256 thread totals
        for (uint32_t off = 1; off < 256; off <<= 1) {
            Fr x = threadIdx.x >= off ? sh[threadIdx.x - off] : op_identity<MUL>();
            __syncthreads();
            sh[threadIdx.x] = op_combine<MUL>(sh[threadIdx.x], x);
            __syncthreads();
        }
        Fr pre = op_combine<MUL>(run, threadIdx.x ? sh[threadIdx.x - 1] : op_identity<MUL>());
#pragma unroll 8
        for (uint32_t k = 0; k < 8; k++) {
            if (b0 + k < nb) fe_store<FrParams>(&tot[b0 + k], pre);
            pre = op_combine<MUL>(pre, loc[k]);
        }
        __syncthreads();
        if (threadIdx.x == 255) run = op_combine<MUL>(run, sh[255]);
        __syncthreads();
    }
}

// z[i] = offset_block + exclusive in-block prefix; last_z gets z[last_index]
template <bool MUL>
__global__ __launch_bounds__(256) void k_fscan_apply(Fr *z, uint32_t n, const Fr *__restrict__ tot,
                                                     Fr *last_z, int64_t last_index) {
    __shared__ Fr sh[256];
    uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * 8;
    Fr loc[8];
    Fr acc = op_identity<MUL>();
#pragma unroll 8
    for (uint32_t k = 0; k < 8; k++) {
        loc[k] = base + k < n ? fe_load<FrParams>(&z[base + k]) : op_identity<MUL>();
        acc = op_combine<MUL>(acc, loc[k]);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t off = 1; off < 256; off <<= 1) {
        Fr x = threadIdx.x >= off ? sh[threadIdx.x - off] : op_identity<MUL>();
        __syncthreads();
        sh[threadIdx.x] = op_combine<MUL>(sh[threadIdx.x], x);
        __syncthreads();
    }
    Fr pre = op_combine<MUL>(fe_load<FrParams>(&tot[blockIdx.x]),
                             threadIdx.x ? sh[threadIdx.x - 1] : op_identity<MUL>());
#pragma unroll 8
    for (uint32_t k = 0; k < 8; k++) {
        uint32_t i = base + k;
        if (i < n) {
            fe_store<FrParams>(&z[i], pre);
            if (last_z && (int64_t)i == last_index) fe_store<FrParams>(last_z, pre);
        }
        pre = op_combine<MUL>(pre, loc[k]);
    }
}

}  // namespace

namespace zk {

// scratch Fr needed by batch_invert_ex for n elements: n (the prefix
// products) up to 2^24 elements, beyond that the multi-level hierarchy's
size_t binv_scratch_elems(uint32_t n) {
    if (n <= (size_t)kTopMax * kTile) return n;
    size_t s = 0;
    uint32_t cnt = n;
    while (cnt > kTop) {
        uint32_t T = (cnt + kM - 1) / kM;
        s += 2 * (size_t)T;
        cnt = T;
    }
    return s;
}

// out[i] = 1/(x[i] (+ beta)) (mode OUT_SET) or out[i] += that (OUT_ADD).
// out may alias x.  temp must hold binv_scratch_elems(n) Fr.
hipError_t batch_invert_ex(const Fr *x, const Fr *beta, Fr *out, int mode, uint32_t n, Fr *temp,
                           hipStream_t stream) {
    if (n == 0) return hipSuccess;
    if (n <= (size_t)kTopMax * kTile) {
        const uint32_t T = (n + kTile - 1) / kTile;
        k_binv_up<<<T, kBT, 0, stream>>>(x, beta, temp, n);
        ZK_CHECK(hipGetLastError());
        k_binv_tiles<<<(T + 63) / 64, 64, 0, stream>>>(temp, T);
        ZK_CHECK(hipGetLastError());
        if (mode == OUT_ADD)
            k_binv_down<OUT_ADD><<<T, kBT, 0, stream>>>(x, beta, temp, out, n);
        else
            k_binv_down<OUT_SET><<<T, kBT, 0, stream>>>(x, beta, temp, out, n);
        return hipGetLastError();
    }
    if (n <= kTop) {
        if (mode == OUT_ADD)
            k_binv_top<OUT_ADD><<<1, 256, 0, stream>>>(x, beta, out, n);
        else
            k_binv_top<OUT_SET><<<1, 256, 0, stream>>>(x, beta, out, n);
        return hipGetLastError();
    }
    const Fr *vals[16];
    const Fr *betas[16];
    Fr *tots[16], *invs[16];
    uint32_t cnts[16], Ts[16];
    int levels = 0;
    const Fr *cur = x;
    const Fr *cur_beta = beta;
    uint32_t cnt = n;
    Fr *tp = temp;
    while (cnt > kTop && levels < 15) {
        uint32_t T = (cnt + kM - 1) / kM;
        vals[levels] = cur;
        betas[levels] = cur_beta;
        cnts[levels] = cnt;
        Ts[levels] = T;
        tots[levels] = tp;
        invs[levels] = tp + T;
        tp += 2 * (size_t)T;
        k_binv_totals<<<(T + 255) / 256, 256, 0, stream>>>(cur, cur_beta, tots[levels], cnt, T);
        ZK_CHECK(hipGetLastError());
        cur = tots[levels];
        cur_beta = nullptr;
        cnt = T;
        levels++;
    }
    if (cnt > kTop) return hipErrorInvalidValue;
    k_binv_top<OUT_SET><<<1, 256, 0, stream>>>(tots[levels - 1], nullptr, invs[levels - 1], cnt);
    ZK_CHECK(hipGetLastError());
    for (int L = levels - 1; L >= 0; L--) {
        uint32_t T = Ts[L];
        uint32_t blocks = (T + 255) / 256;
        if (L > 0)
            k_binv_apply<OUT_SET><<<blocks, 256, 0, stream>>>(vals[L], nullptr, invs[L], invs[L - 1], cnts[L], T);
        else if (mode == OUT_ADD)
            k_binv_apply<OUT_ADD><<<blocks, 256, 0, stream>>>(vals[0], betas[0], invs[0], out, cnts[0], T);
        else
            k_binv_apply<OUT_SET><<<blocks, 256, 0, stream>>>(vals[0], betas[0], invs[0], out, cnts[0], T);
        ZK_CHECK(hipGetLastError());
    }
    return hipSuccess;
}

size_t fscan_scratch_elems(uint32_t n) { return (n + kScanBlock - 1) / kScanBlock + 1; }

// exclusive scan in place; carry (device) = start value or null for the
// identity; last_z (device, may be null) receives z[last_index].
hipError_t field_scan(Fr *z, uint32_t n, bool mul, const Fr *carry, Fr *last_z, int64_t last_index, Fr *tot,
                      hipStream_t stream) {
    if (n == 0) return hipSuccess;
    uint32_t nb = (n + kScanBlock - 1) / kScanBlock;
    if (mul) {
        k_fscan_totals<true><<<nb, 256, 0, stream>>>(z, n, tot);
        ZK_CHECK(hipGetLastError());
        k_fscan_top<true><<<1, 256, 0, stream>>>(tot, nb, carry);
        ZK_CHECK(hipGetLastError());
        k_fscan_apply<true><<<nb, 256, 0, stream>>>(z, n, tot, last_z, last_index);
    } else {
        k_fscan_totals<false><<<nb, 256, 0, stream>>>(z, n, tot);
        ZK_CHECK(hipGetLastError());
        k_fscan_top<false><<<1, 256, 0, stream>>>(tot, nb, carry);
        ZK_CHECK(hipGetLastError());
        k_fscan_apply<false><<<nb, 256, 0, stream>>>(z, n, tot, last_z, last_index);
    }
    return hipGetLastError();
}

}  // namespace zk

namespace zk {

// out = (x + beta)^-1 element-wise (zeros stay zero); temp: n Fr
hipError_t batch_invert_shifted(void *out, const void *x, const void *beta, void *temp, int n, hipStream_t stream) {
    if (n <= 0 || !out || !x || !beta || !temp) return hipErrorInvalidValue;
    return batch_invert_ex((const Fr *)x, (const Fr *)beta, (Fr *)out, OUT_SET, (uint32_t)n, (Fr *)temp, stream);
}

// eval_logup_z_compose after its table inversion: z from the logup sum
// `input` (clobbered as scan scratch), the inverted table (t + beta)^-1 and
// the multiplicities; the same steps and bytes as the drop-in
hipError_t logup_z_inverted(void *z, void *input, const void *table_inv, const void *multiplicity, void *last_z,
                            int last_z_index, int n, hipStream_t stream) {
    if (n <= 0 || !z || !input || !table_inv || !multiplicity || !last_z || last_z_index < 0 || last_z_index >= n)
        return hipErrorInvalidValue;
    uint32_t blocks = ((uint32_t)n + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    k_logup_step2<<<blocks, 256, 0, stream>>>((Fr *)z, (const Fr *)input, (const Fr *)table_inv,
                                               (const Fr *)multiplicity, (uint32_t)n);
    ZK_CHECK(hipGetLastError());
    return field_scan((Fr *)z, (uint32_t)n, false, (const Fr *)last_z, (Fr *)last_z, last_z_index, (Fr *)input,
                      stream);
}

}  // namespace zk

extern "C" {

// In-place batch inversion; temp holds n Fr.
hipError_t zk_batch_invert(void *a, void *temp, int n, hipStream_t stream) {
    if (n < 0 || (n > 0 && !a) || (binv_scratch_elems((uint32_t)n) && !temp)) return hipErrorInvalidValue;
    return batch_invert_ex((const Fr *)a, nullptr, (Fr *)a, OUT_SET, (uint32_t)n, (Fr *)temp, stream);
}

// cuda/bn254.cu:1914-1934: sum[i] (=, or += when init != 0) 1/(input[i] + beta).
// temp holds n Fr.
hipError_t logup_sum_input_inv(void *sum, const void *input, void *temp, const void *beta, int init, int n,
                               hipStream_t stream) {
    if (n <= 0 || !sum || !input || !beta || (binv_scratch_elems((uint32_t)n) && !temp)) return hipErrorInvalidValue;
    return batch_invert_ex((const Fr *)input, (const Fr *)beta, (Fr *)sum, init ? OUT_ADD : OUT_SET, (uint32_t)n,
                           (Fr *)temp, stream);
}

// cuda/bn254.cu:1936-1966: z in place; input (n Fr) is scratch for the scan.
hipError_t eval_logup_z_pure(void *z, void *input, void *table, void *last_z, int last_z_index, int n,
                             hipStream_t stream) {
    (void)table;
    if (n <= 0 || !z || !input || !last_z || last_z_index < 0 || last_z_index >= n) return hipErrorInvalidValue;
    return field_scan((Fr *)z, (uint32_t)n, false, (const Fr *)last_z, (Fr *)last_z, last_z_index, (Fr *)input,
                      stream);
}

// cuda/bn254.cu:1968-1993.  As in the reference, table is inverted in place
// (1/(table + beta)) with z as temporary storage, then
// z[i] = input[i] - m[i] * table[i], then the exclusive scan with carry
// last_z[0] using input as scratch; last_z[0] = z[last_z_index] at the end.
hipError_t eval_logup_z_compose(void *z, void *input, void *table, const void *multiplicity, const void *beta,
                                void *last_z, int last_z_index, int n, hipStream_t stream) {
    if (n <= 0 || !z || !input || !table || !multiplicity || !beta || !last_z || last_z_index < 0 ||
        last_z_index >= n)
        return hipErrorInvalidValue;
    ZK_CHECK(batch_invert_ex((const Fr *)table, (const Fr *)beta, (Fr *)table, OUT_SET, (uint32_t)n, (Fr *)z,
                             stream));
    uint32_t blocks = ((uint32_t)n + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    k_logup_step2<<<blocks, 256, 0, stream>>>((Fr *)z, (const Fr *)input, (const Fr *)table, (const Fr *)multiplicity,
                                               (uint32_t)n);
    ZK_CHECK(hipGetLastError());
    return field_scan((Fr *)z, (uint32_t)n, false, (const Fr *)last_z, (Fr *)last_z, last_z_index, (Fr *)input,
                      stream);
}

}  // extern "C"
