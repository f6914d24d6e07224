// Multi-open arithmetic for gfx950: polynomial evaluation, division by
// (X - z), and the Shplonk h(X) kernels.
//
// Drop-in for `poly_eval` (cuda/bn254.cu:238-256, :1608-1633; caller
// src/eval_poly.rs:18-140), `shplonk_h_x_merge` and `shplonk_h_x_div_points`
// (cuda/bn254.cu:1133-1170, :1580-1606; caller src/multiopen.rs:380-448),
// plus zk_kate_divide: the division by (X - z) that GWC and Shplonk do on the
// host (src/multiopen.rs:145-159, :543-553), here as a device scan.
//
// Both evaluation and division are the weighted suffix scan
//     q_i = b_i + z * q_{i+1},   q_n = 0
// (b = a for evaluation, whose answer is q_0 = a(z); b_i = a_{i+1} for
// division, whose quotient is q_0..q_{n-2}).  It runs as a three-kernel block
// scan.  Block of 256 threads x 8 consecutive elements: each thread folds its
// elements with z, a Hillis-Steele scan over the threads with factors
// z^(8*2^s) (kept in LDS) combines them, and block b's value at its first
// element is its total.  One workgroup then scans the block totals from the
// end with factor z^2048, which gives every block the value q at its end
// (its carry) and the grand total.  The apply pass redoes the block scan with
// the carry folded in and writes q.  Work is one multiplication per element
// per pass; traffic two reads and one write per element.
//
// div_points inverts prod_j (w^i - z_j) for every row with a per-workgroup
// Montgomery trick: 4 rows per thread (strided by 256, so loads coalesce),
// a product tree over the 256 thread products in LDS, one binary-Euclid
// inverse per workgroup, then prefix products in registers to split each
// thread's inverse over its rows.
#include <string.h>

#include "ff29.h"
#include "zkcore.h"

using namespace zk;

namespace {

constexpr uint32_t kT = 256;
constexpr uint32_t kE = 8;
constexpr uint32_t kBlk = kT * kE;  // elements per scan block
constexpr int kLogT = 8;
constexpr uint32_t kDE = 4;           // rows per thread in div_points
constexpr uint32_t kDBlk = kT * kDE;

__host__ __device__ inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// pw[s] = z^(8 * 2^s) for s < kLogT, computed by thread 0 (a barrier must
// follow before the other threads read pw)
__device__ __forceinline__ void scan_powers(const Fr &z, Fr *pw) {
    if (threadIdx.x == 0) {
        Fr p = z * z;
        p = p * p;
        p = p * p;  // z^8
        for (int s = 0; s < kLogT; s++) {
            pw[s] = p;
            p = p * p;
        }
    }
}

// inclusive suffix scan over the block's threads: returns
// S_t = sum_{t' >= t} T_{t'} z^(8 (t' - t)); S is LDS of kT + 1 entries and
// holds S_t at [t] (and 0 at [kT]) afterwards
__device__ __forceinline__ Fr block_suffix(Fr T, Fr *S, const Fr *pw) {
    const uint32_t t = threadIdx.x;
    S[t] = T;
    if (t == 0) S[kT] = fe_zero<FrParams>();
    __syncthreads();
#pragma unroll 1
    for (int s = 0; s < kLogT; s++) {
        const uint32_t d = 1u << s;
        Fr o = t + d < kT ? S[t + d] : fe_zero<FrParams>();
        __syncthreads();
        if (t + d < kT) T = T + pw[s] * o;
        S[t] = T;
        __syncthreads();
    }
    return T;
}

// b_i = src[i + shift] for i + shift < n, else 0; the element at the next
// block's first index comes from `next` (the apply pass writes in place)
__device__ __forceinline__ Fr load_b(const Fr *src, uint64_t i, uint32_t n, uint32_t shift, uint64_t blk_end,
                                     const Fr *next) {
    const uint64_t j = i + shift;
    if (j >= n) return fe_zero<FrParams>();
    if (next && j >= blk_end) return fe_load<FrParams>(next);
    return fe_load<FrParams>(&src[j]);
}

// block totals: tot[b] = q at the block's first element with no carry; also
// first[b] = src[b * kBlk] (for the in-place apply)
__global__ __launch_bounds__(kT) void k_sfx_totals(const Fr *__restrict__ src, uint32_t n, uint32_t shift,
                                                   const Fr *__restrict__ zp, Fr *__restrict__ tot,
                                                   Fr *__restrict__ first) {
    __shared__ Fr S[kT + 1];
    __shared__ Fr pw[kLogT];
    const Fr z = fe_load<FrParams>(zp);
    scan_powers(z, pw);
    const uint64_t base = (uint64_t)blockIdx.x * kBlk;
    const uint64_t t0 = base + (uint64_t)threadIdx.x * kE;
    Fr L = fe_zero<FrParams>();
#pragma unroll
    for (int e = kE - 1; e >= 0; e--) L = load_b(src, t0 + e, n, shift, ~0ull, nullptr) + z * L;
    __syncthreads();  // pw
    Fr Sv = block_suffix(L, S, pw);
    if (threadIdx.x == 0) {
        fe_store<FrParams>(&tot[blockIdx.x], Sv);
        if (first) fe_store<FrParams>(&first[blockIdx.x], fe_load<FrParams>(&src[base]));
    }
}

// one workgroup: carry[b] = sum_{b' > b} tot[b'] W^(b' - b - 1) with
// W = z^kBlk, and *total = sum_b tot[b] W^b (either may be null)
__global__ __launch_bounds__(kT) void k_sfx_top(const Fr *tot, uint32_t nb, const Fr *__restrict__ zp, Fr *carry,
                                                Fr *total) {
    __shared__ Fr S[kT + 1];
    __shared__ Fr pw[kLogT];
    __shared__ Fr cin;
    const uint32_t t = threadIdx.x;
    Fr W = fe_load<FrParams>(zp);
    for (uint32_t s = 1; s < kBlk; s <<= 1) W = W * W;  // z^2048
    scan_powers(W, pw);
    if (t == 0) cin = fe_zero<FrParams>();
    __syncthreads();
    const uint32_t nchunks = ceil_div(nb, kBlk);
    for (uint32_t c = nchunks; c-- > 0;) {
        const uint32_t t0 = c * kBlk + t * kE;
        Fr v[kE];
#pragma unroll
        for (uint32_t e = 0; e < kE; e++) v[e] = t0 + e < nb ? fe_load<FrParams>(&tot[t0 + e]) : fe_zero<FrParams>();
        const Fr carry_in = cin;
        Fr L = t == kT - 1 ? carry_in : fe_zero<FrParams>();
#pragma unroll
        for (int e = kE - 1; e >= 0; e--) L = v[e] + W * L;
        Fr Sv = block_suffix(L, S, pw);
        Fr cur = t == kT - 1 ? carry_in : S[t + 1];
#pragma unroll
        for (int e = kE - 1; e >= 0; e--) {
            if (carry && t0 + e < nb) fe_store<FrParams>(&carry[t0 + e], cur);
            cur = v[e] + W * cur;
        }
        __syncthreads();  // every read of cin and S is done
        if (t == 0) cin = Sv;
        __syncthreads();
    }
    if (t == 0 && total) fe_store<FrParams>(total, cin);
}

// q_i = b_i + z q_{i+1} with the block's carry; dst may equal src.  With
// eval non-null, block 0 also writes a(z) = first[0] + z q_0 (division only)
__global__ __launch_bounds__(kT) void k_sfx_apply(Fr *dst, const Fr *src, uint32_t n, uint32_t shift,
                                                  const Fr *__restrict__ zp, const Fr *__restrict__ carry,
                                                  const Fr *__restrict__ first, Fr *eval) {
    __shared__ Fr S[kT + 1];
    __shared__ Fr pw[kLogT];
    const uint32_t t = threadIdx.x, b = blockIdx.x;
    const Fr z = fe_load<FrParams>(zp);
    scan_powers(z, pw);
    const uint64_t base = (uint64_t)b * kBlk, blk_end = base + kBlk;
    const uint64_t t0 = base + (uint64_t)t * kE;
    const Fr *next = (blk_end < n) ? &first[b + 1] : nullptr;
    const Fr cb = fe_load<FrParams>(&carry[b]);
    Fr v[kE];
#pragma unroll
    for (uint32_t e = 0; e < kE; e++) v[e] = load_b(src, t0 + e, n, shift, blk_end, next);
    Fr L = t == kT - 1 ? cb : fe_zero<FrParams>();
#pragma unroll
    for (int e = kE - 1; e >= 0; e--) L = v[e] + z * L;
    __syncthreads();  // pw; every load of src is done before any store
    block_suffix(L, S, pw);
    Fr cur = t == kT - 1 ? cb : S[t + 1];
#pragma unroll
    for (int e = kE - 1; e >= 0; e--) {
        cur = v[e] + z * cur;
        if (t0 + e < n) fe_store<FrParams>(&dst[t0 + e], cur);
    }
    if (eval && b == 0 && t == 0) fe_store<FrParams>(eval, fe_load<FrParams>(&first[0]) + z * cur);
}

__global__ __launch_bounds__(kT) void k_h_x_merge(Fr *res, const Fr *__restrict__ v, const Fr *__restrict__ values,
                                                  const Fr *__restrict__ omegas, const Fr *__restrict__ diff,
                                                  int diff_n, uint32_t n) {
    const Fr vv = fe_load<FrParams>(v);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        Fr t = fe_load<FrParams>(&values[i]);
        const Fr w = fe_load<FrParams>(&omegas[i]);
        for (int j = 0; j < diff_n; j++) t = t * (w - fe_load<FrParams>(&diff[j]));
        fe_store<FrParams>(&res[i], fe_load<FrParams>(&res[i]) * vv + t);
    }
}

// values[i] /= prod_j (omegas[i] - points[j]); a zero product gives 0
__device__ __forceinline__ Fr row_den(const Fr *omegas, const Fr *points, int points_n, uint64_t i, uint32_t n) {
    if (i >= n) return fe_one<FrParams>();
    const Fr w = fe_load<FrParams>(&omegas[i]);
    Fr x = w - fe_load<FrParams>(&points[0]);
    for (int j = 1; j < points_n; j++) x = x * (w - fe_load<FrParams>(&points[j]));
    return x;
}

// The denominators are computed twice (product pass, then back-substitution)
// so that nothing but the thread's product is live across the inverse, which
// is a function call.
__global__ __launch_bounds__(kT) void k_h_x_div_points(Fr *values, const Fr *__restrict__ omegas,
                                                       const Fr *__restrict__ points, int points_n, uint32_t n) {
    __shared__ Fr tree[2 * kT];
    const uint32_t t = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * kDBlk;
    Fr acc = fe_one<FrParams>();
    for (uint32_t e = 0; e < kDE; e++) {
        const Fr x = row_den(omegas, points, points_n, base + t + (uint64_t)e * kT, n);
        if (!fe_is_zero<FrParams>(x)) acc = acc * x;
    }
    tree[kT + t] = acc;
    __syncthreads();
    for (uint32_t w = kT / 2; w >= 1; w >>= 1) {
        if (t < w) tree[w + t] = tree[2 * (w + t)] * tree[2 * (w + t) + 1];
        __syncthreads();
    }
    if (t == 0) tree[1] = fe_inv<FrParams>(tree[1]);
    __syncthreads();
    for (uint32_t w = 1; w < kT; w <<= 1) {
        if (t < w) {
            const uint32_t p = w + t;
            const Fr ip = tree[p], l = tree[2 * p], r = tree[2 * p + 1];
            tree[2 * p] = ip * r;
            tree[2 * p + 1] = ip * l;
        }
        __syncthreads();
    }
    // points outermost, rows innermost and unrolled: d[] and pre[] stay in
    // registers
    Fr d[kDE], pre[kDE];
    {
        Fr w[kDE];
        const Fr p0 = fe_load<FrParams>(&points[0]);
#pragma unroll
        for (uint32_t e = 0; e < kDE; e++) {
            const uint64_t i = base + t + (uint64_t)e * kT;
            w[e] = i < n ? fe_load<FrParams>(&omegas[i]) : fe_zero<FrParams>();
            d[e] = w[e] - p0;
        }
        for (int j = 1; j < points_n; j++) {
            const Fr pj = fe_load<FrParams>(&points[j]);
#pragma unroll
            for (uint32_t e = 0; e < kDE; e++) d[e] = d[e] * (w[e] - pj);
        }
    }
    acc = fe_one<FrParams>();
#pragma unroll
    for (uint32_t e = 0; e < kDE; e++) {
        if (base + t + (uint64_t)e * kT >= n) d[e] = fe_one<FrParams>();
        pre[e] = acc;
        if (!fe_is_zero<FrParams>(d[e])) acc = acc * d[e];
    }
    Fr inv = tree[kT + t];
#pragma unroll
    for (int e = kDE - 1; e >= 0; e--) {
        const uint64_t i = base + t + (uint64_t)e * kT;
        Fr r = fe_zero<FrParams>();
        if (!fe_is_zero<FrParams>(d[e])) {
            r = inv * pre[e];
            inv = inv * d[e];
        }
        if (i < n) fe_store<FrParams>(&values[i], fe_load<FrParams>(&values[i]) * r);
    }
}

// ---- many polynomials at one point ---------------------------------------
// p_m(x) for up to kEvalItems polynomials of n coefficients per launch pair.
// k_peval_pows: pw[k] = x^(2^k), k < 32, by one thread (31 squarings, once
// per point).  k_peval_blocks, block (b, m): thread t folds its 8 coefficients
// b * 2048 + 8t .. + 7 by Horner with x, then a tree over the 256 thread
// values with factors x^(8 * 2^s) leaves the block's chunk value
// sum_j a_(b*2048+j) x^j.  k_peval_combine, one workgroup per polynomial, folds
// the nb chunk values with W = x^2048 the same way (8 per thread, then the
// tree), and writes p_m(x) to the item's output.  One read of every
// coefficient; no carries or suffix values are written.
constexpr int kEvalItems = 64;
struct EvalArgs {
    const Fr *poly[kEvalItems];
    Fr *out[kEvalItems];
};

__global__ void k_peval_pows(const Fr *__restrict__ xp, Fr *__restrict__ pw) {
    if (threadIdx.x != 0) return;
    Fr p = fe_load<FrParams>(xp);
    for (int k = 0; k < 32; k++) {
        fe_store<FrParams>(&pw[k], p);
        p = p * p;
    }
}

// sum_t v_t x^(8t) over the block's 256 thread values (tree in LDS; the
// factor of level s is pw[3 + s] = x^(8 * 2^s) relative to the step e^lvl)
__device__ __forceinline__ Fr block_fold(Fr v, Fr *S, const Fr *pw, int first_pow) {
    const uint32_t t = threadIdx.x;
    S[t] = v;
    __syncthreads();
#pragma unroll 1
    for (int s = 0; s < kLogT; s++) {
        const uint32_t d = 1u << s;
        if ((t & (2 * d - 1)) == 0) S[t] = S[t] + fe_load<FrParams>(&pw[first_pow + s]) * S[t + d];
        __syncthreads();
    }
    return S[0];
}

__global__ __launch_bounds__(kT) void k_peval_blocks(const EvalArgs a, uint32_t n, uint32_t nb,
                                                     const Fr *__restrict__ pw, Fr *__restrict__ tot) {
    __shared__ Fr S[kT];
    const uint32_t b = blockIdx.x, m = blockIdx.y;
    const Fr *__restrict__ src = a.poly[m];
    const F29 x29 = f29_from32(fe_load<FrParams>(&pw[0]));  // x in the 29-bit product's form
    const uint64_t t0 = (uint64_t)b * kBlk + (uint64_t)threadIdx.x * kE;
    Fr L = fe_zero<FrParams>();
#pragma unroll
    for (int e = kE - 1; e >= 0; e--) {
        const Fr c = t0 + e < n ? fe_load<FrParams>(&src[t0 + e]) : fe_zero<FrParams>();
        L = c + fr_mul_w29(L, x29);
    }
    const Fr v = block_fold(L, S, pw, 3);
    if (threadIdx.x == 0) fe_store<FrParams>(&tot[(size_t)m * nb + b], v);
}

__global__ __launch_bounds__(kT) void k_peval_combine(const EvalArgs a, uint32_t nb, const Fr *__restrict__ pw,
                                                      const Fr *__restrict__ tot) {
    __shared__ Fr S[kT];
    const uint32_t m = blockIdx.x;
    const Fr W = fe_load<FrParams>(&pw[11]);  // x^2048
    Fr acc = fe_zero<FrParams>();
    // chunks of kBlk block values: thread t folds values 8t .. 8t + 7 of the
    // chunk, the tree folds the threads (factors W^(8 * 2^s) = x^(2^(14 + s)))
    const uint32_t nchunks = ceil_div(nb, kBlk);
    Fr Wc = fe_one<FrParams>();  // W^(kBlk * c) for chunk c
    for (uint32_t c = 0; c < nchunks; c++) {
        const uint32_t t0 = c * kBlk + threadIdx.x * kE;
        Fr L = fe_zero<FrParams>();
#pragma unroll
        for (int e = kE - 1; e >= 0; e--) {
            const Fr v = t0 + e < nb ? fe_load<FrParams>(&tot[(size_t)m * nb + t0 + e]) : fe_zero<FrParams>();
            L = v + W * L;
        }
        const Fr part = block_fold(L, S, pw, 14);
        acc = acc + Wc * part;
        Wc = Wc * fe_load<FrParams>(&pw[22]);  // W^2048 = x^(2^22)
        __syncthreads();
    }
    if (threadIdx.x == 0) fe_store<FrParams>(a.out[m], acc);
}

// out[j] = out_in[j] v^k + sum_i v^(k-1-i) p_i[j] over the launch's k
// polynomials (Horner per element; out_in = 0 on the first launch): one read
// of every polynomial and one write of out, instead of a read-modify-write
// pass over out per polynomial
__global__ __launch_bounds__(kT) void k_horner_combine(const EvalArgs a, uint32_t k, bool first,
                                                       const Fr *__restrict__ vp, Fr *__restrict__ out, uint32_t n) {
    const F29 v29 = f29_from32(fe_load<FrParams>(vp));  // v in the 29-bit product's form
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        Fr acc = first ? fe_load<FrParams>(&a.poly[0][j])
                       : fr_mul_w29(fe_load<FrParams>(&out[j]), v29) + fe_load<FrParams>(&a.poly[0][j]);
        for (uint32_t i = 1; i < k; i++) acc = fr_mul_w29(acc, v29) + fe_load<FrParams>(&a.poly[i][j]);
        fe_store<FrParams>(&out[j], acc);
    }
}

uint32_t grid_for(uint32_t n) {
    uint32_t b = ceil_div(n, kT);
    return b < 16384 ? (b ? b : 1) : 16384;
}

}  // namespace

namespace zk {

// out = sum_i v^(count-1-i) polys[i] (n elements each, device pointers in a
// host array; out may alias none of them), Horner per element
hipError_t poly_horner_combine(void *out, const void *const *polys, int count, const void *v, int n,
                               hipStream_t stream) {
    if (!out || !polys || count < 1 || !v || n < 1) return hipErrorInvalidValue;
    for (int first = 0; first < count; first += kEvalItems) {
        const int k = count - first < kEvalItems ? count - first : kEvalItems;
        EvalArgs a;
        memset(&a, 0, sizeof(a));
        for (int m = 0; m < k; m++) {
            if (!polys[first + m] || polys[first + m] == out) return hipErrorInvalidValue;
            a.poly[m] = (const Fr *)polys[first + m];
        }
        k_horner_combine<<<grid_for((uint32_t)n), kT, 0, stream>>>(a, (uint32_t)k, first == 0, (const Fr *)v,
                                                                   (Fr *)out, (uint32_t)n);
        ZK_CHECK(hipGetLastError());
    }
    return hipSuccess;
}

}  // namespace zk

extern "C" {

size_t zk_kate_divide_scratch_bytes(int n) {
    if (n < 1) return 0;
    return (size_t)3 * ceil_div((uint32_t)n, kBlk) * sizeof(Fr);
}

hipError_t zk_kate_divide(void *out, const void *in, const void *z, void *eval, int n, void *temp,
                          hipStream_t stream) {
    if (!out || !in || !z || !temp || n < 1) return hipErrorInvalidValue;
    if (n == 1) {  // the reference's loop leaves a single entry as it is
        if (out != in) ZK_CHECK(hipMemcpyAsync(out, in, sizeof(Fr), hipMemcpyDeviceToDevice, stream));
        if (eval) ZK_CHECK(hipMemcpyAsync(eval, in, sizeof(Fr), hipMemcpyDeviceToDevice, stream));
        return hipSuccess;
    }
    const uint32_t nb = ceil_div((uint32_t)n, kBlk);
    Fr *tot = (Fr *)temp, *carry = tot + nb, *first = carry + nb;
    k_sfx_totals<<<nb, kT, 0, stream>>>((const Fr *)in, (uint32_t)n, 1, (const Fr *)z, tot, first);
    ZK_CHECK(hipGetLastError());
    k_sfx_top<<<1, kT, 0, stream>>>(tot, nb, (const Fr *)z, carry, nullptr);
    ZK_CHECK(hipGetLastError());
    k_sfx_apply<<<nb, kT, 0, stream>>>((Fr *)out, (const Fr *)in, (uint32_t)n, 1, (const Fr *)z, carry, first,
                                       (Fr *)eval);
    return hipGetLastError();
}

// res (n Fr) is scratch; res[0] = p(x) with x = x_table[0]
hipError_t poly_eval(const void *p, void *res, const void *x, int n, hipStream_t stream) {
    if (!p || !res || !x || n < 1) return hipErrorInvalidValue;
    const uint32_t nb = ceil_div((uint32_t)n, kBlk);
    Fr *tot = (Fr *)res + ((uint32_t)n - nb);  // nb < n unless n = 1
    k_sfx_totals<<<nb, kT, 0, stream>>>((const Fr *)p, (uint32_t)n, 0, (const Fr *)x, tot, nullptr);
    ZK_CHECK(hipGetLastError());
    k_sfx_top<<<1, kT, 0, stream>>>(tot, nb, (const Fr *)x, nullptr, (Fr *)res);
    return hipGetLastError();
}

// count polynomials (device pointers in the host array polys, n
// coefficients each) at the point x (device Fr); p_m(x) goes to out[m] (device
// pointers, host array).  scratch: zk_poly_eval_batch_scratch_bytes(n, count).
size_t zk_poly_eval_batch_scratch_bytes(int n, int count) {
    if (n < 1 || count < 1) return 0;
    return (32 + (size_t)count * ceil_div((uint32_t)n, kBlk)) * sizeof(Fr);
}

hipError_t zk_poly_eval_batch(const void *const *polys, void *const *out, int count, const void *x, int n,
                              void *scratch, hipStream_t stream) {
    if (!polys || !out || count < 1 || !x || n < 1 || !scratch) return hipErrorInvalidValue;
    for (int m = 0; m < count; m++)
        if (!polys[m] || !out[m]) return hipErrorInvalidValue;
    const uint32_t nb = ceil_div((uint32_t)n, kBlk);
    Fr *pw = (Fr *)scratch, *tot = pw + 32;
    k_peval_pows<<<1, 64, 0, stream>>>((const Fr *)x, pw);
    ZK_CHECK(hipGetLastError());
    for (int first = 0; first < count; first += kEvalItems) {
        const int k = count - first < kEvalItems ? count - first : kEvalItems;
        EvalArgs a;
        memset(&a, 0, sizeof(a));
        for (int m = 0; m < k; m++) {
            a.poly[m] = (const Fr *)polys[first + m];
            a.out[m] = (Fr *)out[first + m];
        }
        k_peval_blocks<<<dim3(nb, (uint32_t)k), kT, 0, stream>>>(a, (uint32_t)n, nb, pw, tot);
        ZK_CHECK(hipGetLastError());
        k_peval_combine<<<(uint32_t)k, kT, 0, stream>>>(a, nb, pw, tot);
        ZK_CHECK(hipGetLastError());
    }
    return hipSuccess;
}

hipError_t zk_shplonk_h_x_merge_on(void *res, const void *v, const void *values, const void *omegas,
                                   const void *diff_points, int diff_points_n, int n, hipStream_t stream) {
    if (!res || !v || !values || !omegas || n < 0 || diff_points_n < 0 || (diff_points_n && !diff_points))
        return hipErrorInvalidValue;
    if (!n) return hipSuccess;
    k_h_x_merge<<<grid_for((uint32_t)n), kT, 0, stream>>>((Fr *)res, (const Fr *)v, (const Fr *)values,
                                                          (const Fr *)omegas, (const Fr *)diff_points,
                                                          diff_points_n, (uint32_t)n);
    return hipGetLastError();
}

hipError_t shplonk_h_x_merge(void *res, const void *v, const void *values, const void *omegas,
                             const void *diff_points, int diff_points_n, int n) {
    return zk_shplonk_h_x_merge_on(res, v, values, omegas, diff_points, diff_points_n, n, 0);
}

hipError_t zk_shplonk_h_x_div_points_on(void *values, const void *omegas, const void *points, int points_n, int n,
                                        hipStream_t stream) {
    if (!values || !omegas || !points || points_n < 1 || n < 0) return hipErrorInvalidValue;
    if (!n) return hipSuccess;
    k_h_x_div_points<<<ceil_div((uint32_t)n, kDBlk), kT, 0, stream>>>((Fr *)values, (const Fr *)omegas,
                                                                    (const Fr *)points, points_n, (uint32_t)n);
    return hipGetLastError();
}

hipError_t shplonk_h_x_div_points(void *values, const void *omegas, const void *points, int points_n, int n) {
    return zk_shplonk_h_x_div_points_on(values, omegas, points, points_n, n, 0);
}

}  // extern "C"
