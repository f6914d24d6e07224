// BN254 G1 multi-scalar multiplication (Pippenger) for gfx950.
//
// Drop-in for `msm` + `batch_msm_collect` (cuda/msm.cu:339-467, host driver
// src/cuda/msm.rs:239-455).  Same buffers and result location: the caller's
// curve slice receives the MSM result as a Jacobian point in its first 96
// bytes after batch_msm_collect.
//
// Pipeline (one MSM):
//   1. digits   one thread per scalar: Montgomery -> canonical, signed c-bit
//               digits d_w in (-2^(c-1), 2^(c-1)], written window-major.
//               Zero digits add nothing.
//   2. sort     two-level counting sort of (bucket, point index | sign):
//               per-tile LDS histograms of the high 8 bucket bits, one global
//               scan, scatter into (window, high bits) groups with LDS ranks,
//               then an LDS counting sort of each group on the low 8 bits that
//               also emits the bucket offsets.  Writes land in short
//               contiguous runs instead of 4-byte random scatters.
//   4. accumulate  threads own equal chunks of the sorted entries (perfect
//               load balance for any digit distribution); runs of one bucket
//               are summed with XYZZ mixed additions; runs cut by a chunk
//               border are left as head/tail partials.
//   5. fixup    partials of buckets that span chunks are merged.
//   6. reduce   S_w = sum_b (b+1) B_{w,b}: per segment of L1 buckets a
//               running sum gives (WS, SUM); then
//               S_w = sum_s WS_s + L1 * sum_bit 2^bit * (sum_{s has bit} SUM_s),
//               the masked sums being parallel tree reductions.
//   7. collect  sum_w 2^(c w) S_w by Horner over windows, normalised to z = 1
//               so the bytes are canonical.
// The bucket semantics restate cuda/msm.cu:9-181 (digit d of window w adds
// P_i to bucket (w, d)); signed digits halve the buckets per window.
#include <string.h>

#include "zkcore.h"

using namespace zk;

namespace {

constexpr uint32_t kL1 = 8;          // buckets per first-level reduction segment
constexpr uint32_t kChunk = 48;      // sorted entries per accumulation thread
constexpr uint32_t kNone = 0xffffffffu;
constexpr uint32_t kMagic = 0x4d534d31u;  // "MSM1": window sums present in the slice

// Arrays placed into caller memory, in two lifetimes: phase 1 (the sort) and
// phase 2 (accumulation and reduction).  offsets and entries live across both;
// the phase-1 temporaries (tmp, hist, starts, scan) are free again in phase 2.
enum MsmArr {
    A_OFFSETS, A_ENTRIES,                                          // both phases
    A_TMP, A_HIST, A_STARTS, A_SCAN,                               // phase 1
    A_BUCKETS, A_PHEAD, A_PTAIL, A_HB, A_TB, A_WS1, A_SUM1, A_Q,   // phase 2
    A_COUNT
};
constexpr int kFirstPhase2 = A_BUCKETS;

constexpr uint32_t kSortTile = 4096;  // scalars per histogram / scatter tile

struct MsmPlan {
    uint32_t n, c, W, B, NB, T1, nbits;
    uint32_t hi_bits, lo_bits, nbins, ntiles, nhist;
    uint64_t max_entries, chunks;
    size_t size[A_COUNT];
    uint8_t *ptr[A_COUNT];
};

__host__ __device__ inline uint64_t align256(uint64_t x) { return (x + 255) & ~(uint64_t)255; }

MsmPlan make_plan(uint32_t n, uint32_t c) {
    MsmPlan p;
    memset(&p, 0, sizeof(p));
    p.n = n;
    p.c = c;
    p.W = (255 + c - 1) / c;
    p.B = 1u << (c - 1);
    p.NB = p.W * p.B;
    p.T1 = (p.B + kL1 - 1) / kL1;
    p.nbits = log2_exact(p.T1);
    p.max_entries = (uint64_t)n * p.W;
    p.chunks = (p.max_entries + kChunk - 1) / kChunk;
    p.hi_bits = c - 1 > 8 ? c - 1 - 8 : 0;
    p.lo_bits = (c - 1) - p.hi_bits;
    p.nbins = 1u << p.hi_bits;
    p.ntiles = (n + kSortTile - 1) / kSortTile;
    p.nhist = p.W * p.nbins * p.ntiles;
    p.size[A_OFFSETS] = (size_t)(p.NB + 1) * 4;
    // digits (window-major, n * W u32) share this array with the final entries
    p.size[A_ENTRIES] = (size_t)p.max_entries * 4;
    p.size[A_TMP] = (size_t)p.max_entries * 8;
    p.size[A_HIST] = (size_t)p.nhist * 4;
    p.size[A_STARTS] = (size_t)(p.nhist + 1) * 4;
    p.size[A_SCAN] = ((size_t)(p.nhist + 1 + 1023) / 1024 + 1) * 4;
    p.size[A_BUCKETS] = (size_t)p.NB * sizeof(G1Xyzz);
    p.size[A_PHEAD] = (size_t)p.chunks * sizeof(G1Xyzz);
    p.size[A_PTAIL] = (size_t)p.chunks * sizeof(G1Xyzz);
    p.size[A_HB] = (size_t)p.chunks * 4;
    p.size[A_TB] = (size_t)p.chunks * 4;
    p.size[A_WS1] = (size_t)p.W * p.T1 * sizeof(G1Xyzz);
    p.size[A_SUM1] = (size_t)p.W * p.T1 * sizeof(G1Xyzz);
    p.size[A_Q] = (size_t)p.W * (p.nbits + 1) * sizeof(G1Xyzz);
    return p;
}

// bytes that always suffice in one region: every array gets its own block
// (placement may still reuse phase-1 blocks, which only leaves room unused)
size_t plan_total(const MsmPlan &p) {
    size_t t = 0;
    for (int a = 0; a < A_COUNT; a++) t += align256(p.size[a]);
    return t + 256;
}

struct Seg {
    uint8_t *ptr;
    size_t left;
};

// Largest array first, each into the free segment with the most room.
bool place_range(MsmPlan &p, int lo, int hi, Seg *segs, int &nseg) {
    bool done[A_COUNT] = {false};
    for (int k = lo; k < hi; k++) {
        int best = -1;
        for (int a = lo; a < hi; a++)
            if (!done[a] && (best < 0 || p.size[a] > p.size[best])) best = a;
        size_t need = align256(p.size[best]);
        int rr = -1;
        for (int r = 0; r < nseg; r++)
            if (segs[r].left >= need && (rr < 0 || segs[r].left > segs[rr].left)) rr = r;
        if (rr < 0) return false;
        p.ptr[best] = segs[rr].ptr;
        segs[rr].ptr += need;
        segs[rr].left -= need;
        done[best] = true;
    }
    return true;
}

// Places the plan into caller regions (bases aligned up to 256 B).  Phase-2
// arrays may reuse the blocks of the phase-1 temporaries.
bool place(MsmPlan &p, uint8_t *const *base, const size_t *bytes, int nreg) {
    Seg segs[16];
    int nseg = 0;
    for (int r = 0; r < nreg && r < 8; r++) {
        uintptr_t b = (uintptr_t)base[r];
        uintptr_t a = (b + 255) & ~(uintptr_t)255;
        segs[nseg].ptr = (uint8_t *)a;
        segs[nseg].left = (base[r] && bytes[r] > a - b) ? bytes[r] - (a - b) : 0;
        nseg++;
    }
    if (!place_range(p, A_OFFSETS, kFirstPhase2, segs, nseg)) return false;
    for (int a = A_TMP; a < kFirstPhase2; a++) {  // phase-1 temporaries become free
        segs[nseg].ptr = p.ptr[a];
        segs[nseg].left = align256(p.size[a]);
        nseg++;
    }
    return place_range(p, kFirstPhase2, A_COUNT, segs, nseg);
}

// Default window for a single MSM of n points (the autotune table).
uint32_t default_window(uint32_t n) {
    uint32_t lg = log2_exact(n ? n : 1);
    int c = (int)lg - 5;
    if (c < 6) c = 6;
    if (c > 17) c = 17;
    return (uint32_t)c;
}

__device__ __forceinline__ uint32_t scalar_bits(const uint32_t s[8], uint32_t start, uint32_t c) {
    uint32_t limb = start >> 5, off = start & 31;
    uint32_t v = s[limb] >> off;
    if (off && limb + 1 < 8) v |= s[limb + 1] << (32 - off);
    return v & ((1u << c) - 1);
}

// Signed digit of window w given the running carry; returns d in
// (-2^(c-1), 2^(c-1)] and updates carry.
__device__ __forceinline__ int32_t signed_digit(const uint32_t s[8], uint32_t w, uint32_t c, uint32_t &carry) {
    uint32_t start = w * c;
    uint32_t v = (start < 256 ? scalar_bits(s, start, c) : 0u) + carry;
    uint32_t half = 1u << (c - 1);
    if (v > half) {
        carry = 1;
        return (int32_t)v - (int32_t)(1u << c);
    }
    carry = 0;
    return (int32_t)v;
}

// digits[w * n + i] = 0 for a zero digit, else |d| | (d < 0) << 31 with
// |d| in [1, 2^(c-1)] (bucket |d| - 1)
__global__ __launch_bounds__(256) void k_msm_digits(const Fr *__restrict__ scalars, uint32_t n, uint32_t c,
                                                    uint32_t W, uint32_t *__restrict__ digits) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Fr s = fe_from_mont<FrParams>(fe_load<FrParams>(&scalars[i]));
    uint32_t carry = 0;
    for (uint32_t w = 0; w < W; w++) {
        int32_t d = signed_digit(s.v, w, c, carry);
        digits[(size_t)w * n + i] = d ? ((uint32_t)(d < 0 ? -d : d) | (d < 0 ? 0x80000000u : 0u)) : 0u;
    }
}

// per (tile, window): histogram of the high bucket bits
// hist[(w * nbins + bin) * ntiles + tile]
__global__ __launch_bounds__(256) void k_msm_hist(const uint32_t *__restrict__ digits, uint32_t n, uint32_t lo_bits,
                                                  uint32_t nbins, uint32_t ntiles, uint32_t *__restrict__ hist) {
    __shared__ uint32_t h[256];
    const uint32_t tile = blockIdx.x, w = blockIdx.y;
    for (uint32_t k = threadIdx.x; k < nbins; k += blockDim.x) h[k] = 0;
    __syncthreads();
    const uint32_t *dw = digits + (size_t)w * n;
    uint32_t end = (tile + 1) * kSortTile < n ? (tile + 1) * kSortTile : n;
    for (uint32_t i = tile * kSortTile + threadIdx.x; i < end; i += blockDim.x) {
        uint32_t d = dw[i];
        if (d) atomicAdd(&h[((d & 0x7fffffffu) - 1) >> lo_bits], 1u);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nbins; k += blockDim.x) hist[((size_t)w * nbins + k) * ntiles + tile] = h[k];
}

// per (tile, window): place each nonzero digit into its (window, bin) group
// at starts[...] + rank; tmp entry = (low bucket bits, index | sign)
__global__ __launch_bounds__(256) void k_msm_scatter_hi(const uint32_t *__restrict__ digits, uint32_t n,
                                                        uint32_t lo_bits, uint32_t nbins, uint32_t ntiles,
                                                        const uint32_t *__restrict__ starts, uint2 *__restrict__ tmp) {
    __shared__ uint32_t cur[256];
    const uint32_t tile = blockIdx.x, w = blockIdx.y;
    for (uint32_t k = threadIdx.x; k < nbins; k += blockDim.x) cur[k] = starts[((size_t)w * nbins + k) * ntiles + tile];
    __syncthreads();
    const uint32_t *dw = digits + (size_t)w * n;
    const uint32_t lo_mask = (1u << lo_bits) - 1;
    uint32_t end = (tile + 1) * kSortTile < n ? (tile + 1) * kSortTile : n;
    for (uint32_t i = tile * kSortTile + threadIdx.x; i < end; i += blockDim.x) {
        uint32_t d = dw[i];
        if (d) {
            uint32_t b = (d & 0x7fffffffu) - 1;
            uint32_t pos = atomicAdd(&cur[b >> lo_bits], 1u);
            tmp[pos] = make_uint2(b & lo_mask, i | (d & 0x80000000u));
        }
    }
}

// per (window, bin) group: counting sort on the low bucket bits in LDS;
// writes the final entries and the offsets of the group's buckets
__global__ __launch_bounds__(256) void k_msm_sort_lo(const uint2 *__restrict__ tmp, const uint32_t *__restrict__ starts,
                                                     uint32_t nbins, uint32_t ntiles, uint32_t lo_bits, uint32_t B,
                                                     uint32_t ngroups, uint32_t *__restrict__ entries,
                                                     uint32_t *__restrict__ offsets) {
    __shared__ uint32_t h[256];
    __shared__ uint32_t cur[256];
    const uint32_t g = blockIdx.x;
    const uint32_t w = g / nbins, bin = g % nbins;
    const uint32_t nlo = 1u << lo_bits;
    const uint32_t gs = starts[(size_t)g * ntiles];
    const uint32_t ge = starts[(size_t)(g + 1) * ntiles];  // starts has a final total entry
    for (uint32_t k = threadIdx.x; k < nlo; k += blockDim.x) h[k] = 0;
    __syncthreads();
    for (uint32_t p = gs + threadIdx.x; p < ge; p += blockDim.x) atomicAdd(&h[tmp[p].x], 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = gs;
        for (uint32_t k = 0; k < nlo; k++) {
            cur[k] = run;
            run += h[k];
        }
    }
    __syncthreads();
    const uint32_t base = w * B + (bin << lo_bits);
    for (uint32_t k = threadIdx.x; k < nlo; k += blockDim.x) offsets[base + k] = cur[k];
    if (g == ngroups - 1 && threadIdx.x == 0) offsets[(size_t)w * B + B] = ge;  // = total: offsets[NB]
    __syncthreads();
    for (uint32_t p = gs + threadIdx.x; p < ge; p += blockDim.x) {
        uint2 t = tmp[p];
        uint32_t pos = atomicAdd(&cur[t.x], 1u);
        entries[pos] = t.y;
    }
}

// ---- exclusive scan of u32 counts (3 kernels) ---------------------------
__global__ void k_scan_block_sums(const uint32_t *__restrict__ in, uint32_t n, uint32_t *__restrict__ sums) {
    __shared__ uint32_t sh[256];
    uint32_t base = blockIdx.x * 1024;
    uint32_t t = 0;
    for (uint32_t k = threadIdx.x; k < 1024; k += 256) {
        uint32_t i = base + k;
        t += i < n ? in[i] : 0u;
    }
    sh[threadIdx.x] = t;
    __syncthreads();
    for (uint32_t s = 128; s; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = sh[0];
}
// single block: exclusive scan of the block sums, in place
__global__ void k_scan_sums(uint32_t *sums, uint32_t nb) {
    __shared__ uint32_t sh[1024];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += 1024) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < nb ? sums[i] : 0u;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t off = 1; off < 1024; off <<= 1) {
            uint32_t x = threadIdx.x >= off ? sh[threadIdx.x - off] : 0u;
            __syncthreads();
            sh[threadIdx.x] += x;
            __syncthreads();
        }
        if (i < nb) sums[i] = carry + sh[threadIdx.x] - v;
        uint32_t total = sh[1023];
        __syncthreads();
        carry += total;
    }
}
// per block: exclusive scan of 1024 values plus the block offset; writes out
// (offsets) and, if cursor != null, a copy into cursor.  out[n] = total is
// written by the block that owns index n (in has n entries, out n + 1).
__global__ void k_scan_apply(const uint32_t *__restrict__ in, uint32_t n, const uint32_t *__restrict__ sums,
                             uint32_t *__restrict__ out, uint32_t *__restrict__ cursor) {
    __shared__ uint32_t sh[1024];
    uint32_t base = blockIdx.x * 1024;
    uint32_t i = base + threadIdx.x;
    uint32_t v = i < n ? in[i] : 0u;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        uint32_t x = threadIdx.x >= off ? sh[threadIdx.x - off] : 0u;
        __syncthreads();
        sh[threadIdx.x] += x;
        __syncthreads();
    }
    uint32_t excl = sums[blockIdx.x] + sh[threadIdx.x] - v;
    if (i < n) {
        out[i] = excl;
        if (cursor) cursor[i] = excl;
    }
    if (i == n) out[n] = excl;  // v = 0 here, so excl is the grand total
}

// ---- accumulation --------------------------------------------------------
__device__ __forceinline__ uint32_t find_bucket(const uint32_t *__restrict__ off, uint32_t NB, uint32_t pos) {
    // last b with off[b] <= pos (off is non-decreasing, off[0] = 0)
    uint32_t lo = 0, hi = NB;  // invariant: off[lo] <= pos, answer < hi
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= pos) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ G1Xyzz xyzz_add_entry(const G1Xyzz &acc, const G1Aff *__restrict__ pts, uint32_t e) {
    G1Aff q = aff_load(&pts[e & 0x7fffffffu]);
    if (aff_is_id(q)) return acc;
    Fp y = (e >> 31) ? fe_neg<FpParams>(q.y) : q.y;
    return xyzz_add_aff(acc, q.x, y);
}

__global__ __launch_bounds__(256) void k_msm_accumulate(const uint32_t *__restrict__ entries,
                                                        const uint32_t *__restrict__ off, uint32_t NB,
                                                        const G1Aff *__restrict__ pts, G1Xyzz *__restrict__ buckets,
                                                        G1Xyzz *__restrict__ phead, G1Xyzz *__restrict__ ptail,
                                                        uint32_t *__restrict__ hb, uint32_t *__restrict__ tb,
                                                        uint64_t nchunks) {
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nchunks) return;
    const uint32_t total = off[NB];
    const uint32_t s = (uint32_t)(t * kChunk);
    uint32_t head_bucket = kNone, tail_bucket = kNone;
    if (s < total) {
        const uint32_t e = s + kChunk < total ? s + kChunk : total;
        uint32_t b = find_bucket(off, NB, s);
        uint32_t bend = off[b + 1];
        bool first = true, before = off[b] < s;
        G1Xyzz acc = xyzz_id();
        // one uniform loop over the chunk (every lane runs e - s steps); the
        // next entry and its point are loaded one step ahead
        uint32_t ecur = entries[s];
        G1Aff qcur = aff_load(&pts[ecur & 0x7fffffffu]);
        for (uint32_t pos = s; pos < e; pos++) {
            const bool more = pos + 1 < e;
            const uint32_t enext = more ? entries[pos + 1] : 0u;
            const G1Aff qnext = more ? aff_load(&pts[enext & 0x7fffffffu]) : qcur;
            if (!aff_is_id(qcur)) {
                Fp y = (ecur >> 31) ? fe_neg<FpParams>(qcur.y) : qcur.y;
                acc = xyzz_add_aff(acc, qcur.x, y);
            }
            if (pos + 1 == bend || !more) {
                // the run of bucket b ends at this step
                if (first && before) {
                    xyzz_store(&phead[t], acc);
                    head_bucket = b;
                } else if (bend > e) {
                    xyzz_store(&ptail[t], acc);
                    tail_bucket = b;
                } else {
                    xyzz_store(&buckets[b], acc);
                }
                first = false;
                acc = xyzz_id();
                if (more) {
                    // next non-empty bucket starting at pos + 1
                    const uint32_t np = pos + 1;
                    uint32_t nb = b + 1, steps = 0;
                    while (off[nb + 1] == np && steps < 16) {
                        nb++;
                        steps++;
                    }
                    if (off[nb + 1] == np) nb = find_bucket(off, NB, np);
                    b = nb;
                    bend = off[b + 1];
                }
            }
            ecur = enext;
            qcur = qnext;
        }
    }
    hb[t] = head_bucket;
    tb[t] = tail_bucket;
}

// merges the partial sums of buckets that span several chunks
__global__ __launch_bounds__(256) void k_msm_fixup(G1Xyzz *__restrict__ buckets, const G1Xyzz *__restrict__ phead,
                            const G1Xyzz *__restrict__ ptail, const uint32_t *__restrict__ hb,
                            const uint32_t *__restrict__ tb, uint64_t nchunks) {
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nchunks) return;
    uint32_t b = tb[t];
    if (b == kNone) return;
    G1Xyzz acc = xyzz_load(&ptail[t]);
    for (uint64_t u = t + 1; u < nchunks && hb[u] == b; u++) acc = xyzz_add(acc, xyzz_load(&phead[u]));
    xyzz_store(&buckets[b], acc);
}

// buckets that some chunk starts inside and no chunk owns fully are
// covered above; a bucket whose whole range lies in one chunk was written
// directly.  Empty buckets are never read (their count is zero).

// ---- reduction -------------------------------------------------------------
// one thread per (window, segment of kL1 buckets)
__global__ __launch_bounds__(64) void k_msm_reduce_l1(const G1Xyzz *__restrict__ buckets, const uint32_t *__restrict__ off,
                                uint32_t W, uint32_t B, uint32_t T1, G1Xyzz *__restrict__ ws1,
                                G1Xyzz *__restrict__ sum1) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= W * T1) return;
    uint32_t w = g / T1, s = g % T1;
    G1Xyzz run = xyzz_id(), acc = xyzz_id();
    uint32_t lo = s * kL1;
    uint32_t hi = lo + kL1 < B ? lo + kL1 : B;
    for (uint32_t b = hi; b-- > lo;) {
        uint32_t idx = w * B + b;
        if (off[idx + 1] != off[idx]) run = xyzz_add(run, xyzz_load(&buckets[idx]));
        acc = xyzz_add(acc, run);
    }
    xyzz_store(&ws1[g], acc);
    xyzz_store(&sum1[g], run);
}

// grid (nbits + 1, W): q < nbits -> sum of sum1[w][s] over s with bit q set;
// q == nbits -> sum of ws1[w][s] over all s.
__global__ __launch_bounds__(256) void k_msm_masked_sums(const G1Xyzz *__restrict__ ws1,
                                                         const G1Xyzz *__restrict__ sum1, uint32_t T1,
                                                         uint32_t nbits, G1Xyzz *__restrict__ q) {
    __shared__ G1Xyzz sh[256];
    uint32_t bit = blockIdx.x, w = blockIdx.y;
    const G1Xyzz *src = (bit == nbits ? ws1 : sum1) + (size_t)w * T1;
    G1Xyzz acc = xyzz_id();
    for (uint32_t s = threadIdx.x; s < T1; s += blockDim.x) {
        if (bit == nbits || ((s >> bit) & 1u)) acc = xyzz_add(acc, xyzz_load(&src[s]));
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t st = 128; st; st >>= 1) {
        if (threadIdx.x < st) sh[threadIdx.x] = xyzz_add(sh[threadIdx.x], sh[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) xyzz_store(&q[(size_t)w * (nbits + 1) + bit], sh[0]);
}

struct SliceHeader {
    uint32_t magic, c, W, pad[5];
};

// one thread per window: S_w = A + L1 * sum_bit 2^bit Q_bit -> slice
__global__ __launch_bounds__(64) void k_msm_window_finalize(const G1Xyzz *__restrict__ q, uint32_t W, uint32_t nbits, uint32_t c,
                                      uint8_t *__restrict__ slice) {
    uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) return;
    const G1Xyzz *qw = q + (size_t)w * (nbits + 1);
    G1Jac acc = xyzz_to_jac(nbits ? xyzz_load(&qw[nbits - 1]) : xyzz_id());
    for (int b = (int)nbits - 2; b >= 0; b--) {
        acc = jac_dbl(acc);
        acc = jac_add(acc, xyzz_to_jac(xyzz_load(&qw[b])));
    }
    for (uint32_t k = 1; k < kL1; k <<= 1) acc = jac_dbl(acc);
    acc = jac_add(acc, xyzz_to_jac(xyzz_load(&qw[nbits])));
    G1Jac *sums = reinterpret_cast<G1Jac *>(slice + sizeof(G1Jac));
    jac_store(&sums[w], acc);
    if (w == 0) {
        SliceHeader *h = reinterpret_cast<SliceHeader *>(slice + 64);
        h->magic = kMagic;
        h->c = c;
        h->W = W;
    }
}

// Horner over windows: result (normalised Jacobian) at slice[0].
__device__ void collect_slice(uint8_t *slice) {
    const SliceHeader *h = reinterpret_cast<const SliceHeader *>(slice + 64);
    uint32_t c = h->c, W = h->W;
    if (h->magic != kMagic || W == 0 || W > 128 || c == 0 || c > 32) {
        // no window sums in this slice (msm() failed or was not run): emit
        // the identity rather than reading garbage
        G1Jac id;
        id.x = fe_zero<FpParams>();
        id.y = fe_one<FpParams>();
        id.z = fe_zero<FpParams>();
        jac_store(reinterpret_cast<G1Jac *>(slice), id);
        return;
    }
    const G1Jac *sums = reinterpret_cast<const G1Jac *>(slice + sizeof(G1Jac));
    G1Jac acc = jac_load(&sums[W - 1]);
    for (int w = (int)W - 2; w >= 0; w--) {
        for (uint32_t k = 0; k < c; k++) acc = jac_dbl(acc);
        acc = jac_add(acc, jac_load(&sums[w]));
    }
    jac_store(reinterpret_cast<G1Jac *>(slice), jac_normalize(acc));
}

__global__ __launch_bounds__(64) void k_msm_collect(uint8_t *const *__restrict__ slices, uint32_t batch) {
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= batch) return;
    collect_slice(slices[m]);
}
__global__ __launch_bounds__(64) void k_msm_collect_one(uint8_t *slice) {
    if (threadIdx.x == 0 && blockIdx.x == 0) collect_slice(slice);
}

// device-side slice header reset so a stale header is never trusted
__global__ void k_slice_clear(uint8_t *slice) {
    if (threadIdx.x == 0) reinterpret_cast<SliceHeader *>(slice + 64)->magic = 0;
}

// Opt-in phase timing: when enabled, msm_windows records a HIP event on the
// launch stream after each phase, waits for the last one and accumulates
// the per-phase durations (ms) and call counts.
enum MsmPhase { PH_COUNT = 0, PH_SCAN, PH_SCATTER, PH_ACCUMULATE, PH_FIXUP, PH_REDUCE, PH_FINALIZE, PH_NUM };
struct PhaseProf {
    bool on = false;
    double ms[PH_NUM] = {0};
    uint64_t calls = 0;
    hipEvent_t ev[PH_NUM + 1] = {};
    bool init = false;
};
PhaseProf g_prof;

hipError_t prof_mark(int k, hipStream_t stream) {
    if (!g_prof.on) return hipSuccess;
    if (!g_prof.init) {
        for (auto &e : g_prof.ev) ZK_CHECK(hipEventCreate(&e));
        g_prof.init = true;
    }
    return hipEventRecord(g_prof.ev[k], stream);
}
hipError_t prof_finish() {
    if (!g_prof.on) return hipSuccess;
    ZK_CHECK(hipEventSynchronize(g_prof.ev[PH_NUM]));
    for (int k = 0; k < PH_NUM; k++) {
        float ms = 0;
        ZK_CHECK(hipEventElapsedTime(&ms, g_prof.ev[k], g_prof.ev[k + 1]));
        g_prof.ms[k] += ms;
    }
    g_prof.calls++;
    return hipSuccess;
}

hipError_t launch_scan(const uint32_t *in, uint32_t n, uint32_t *out, uint32_t *cursor, uint32_t *sums,
                       hipStream_t stream) {
    uint32_t nb = (n + 1 + 1023) / 1024;
    k_scan_block_sums<<<nb, 256, 0, stream>>>(in, n, sums);
    ZK_CHECK(hipGetLastError());
    k_scan_sums<<<1, 1024, 0, stream>>>(sums, nb);
    ZK_CHECK(hipGetLastError());
    k_scan_apply<<<nb, 1024, 0, stream>>>(in, n, sums, out, cursor);
    return hipGetLastError();
}

}  // namespace

namespace zk {

size_t msm_scratch_bytes(uint32_t n, uint32_t c) { return plan_total(make_plan(n, c)); }
uint32_t msm_default_window(uint32_t n) { return default_window(n); }

// Runs one MSM up to its per-window sums, written into `slice` (Jacobian
// sums at slice + 96, header at slice + 64; slice needs 96 * (W + 1) bytes).
// Scratch arrays are placed into the given regions; no allocation happens.
// c == 0 picks the largest window <= the default that fits the regions.
hipError_t msm_windows(uint8_t *slice, size_t slice_bytes, const G1Aff *points, const Fr *scalars, uint32_t n,
                       uint32_t c, uint8_t *const *regions, const size_t *region_bytes, int nreg,
                       hipStream_t stream) {
    if (n == 0 || n > 0x7fffffffu || !slice || !points || !scalars || nreg < 1 || nreg > 8)
        return hipErrorInvalidValue;
    if (c != 0 && (c < 2 || c > 17)) return hipErrorInvalidValue;  // sort kernels: <= 8 + 8 bucket bits
    MsmPlan p;
    bool ok = false;
    const uint32_t c0 = c ? c : default_window(n);
    // the default window first, then larger windows (fewer entries), then smaller
    uint32_t order[32];
    int norder = 0;
    order[norder++] = c0;
    if (!c) {
        for (uint32_t cc = c0 + 1; cc <= 17; cc++) order[norder++] = cc;
        for (uint32_t cc = c0 - 1; cc >= 2; cc--) order[norder++] = cc;
    }
    for (int k = 0; k < norder && !ok; k++) {
        p = make_plan(n, order[k]);
        ok = (size_t)sizeof(G1Jac) * (p.W + 1) <= slice_bytes && place(p, regions, region_bytes, nreg);
    }
    if (!ok) return hipErrorOutOfMemory;
    c = p.c;
    uint32_t *hist = (uint32_t *)p.ptr[A_HIST];
    uint32_t *offsets = (uint32_t *)p.ptr[A_OFFSETS];
    uint32_t *starts = (uint32_t *)p.ptr[A_STARTS];  // dead after k_msm_sort_lo
    uint32_t *entries = (uint32_t *)p.ptr[A_ENTRIES];
    uint32_t *digits = entries;  // dead before the final entries are written
    uint2 *tmp = (uint2 *)p.ptr[A_TMP];  // dead after k_msm_sort_lo
    G1Xyzz *buckets = (G1Xyzz *)p.ptr[A_BUCKETS];
    G1Xyzz *phead = (G1Xyzz *)p.ptr[A_PHEAD];
    G1Xyzz *ptail = (G1Xyzz *)p.ptr[A_PTAIL];
    uint32_t *hb = (uint32_t *)p.ptr[A_HB];
    uint32_t *tb = (uint32_t *)p.ptr[A_TB];
    G1Xyzz *ws1 = (G1Xyzz *)p.ptr[A_WS1];
    G1Xyzz *sum1 = (G1Xyzz *)p.ptr[A_SUM1];
    G1Xyzz *q = (G1Xyzz *)p.ptr[A_Q];
    uint32_t *scan_sums = (uint32_t *)p.ptr[A_SCAN];

    ZK_CHECK(prof_mark(PH_COUNT, stream));
    k_slice_clear<<<1, 1, 0, stream>>>(slice);
    ZK_CHECK(hipGetLastError());
    const uint32_t blocks = (n + 255) / 256;
    k_msm_digits<<<blocks, 256, 0, stream>>>(scalars, n, c, p.W, digits);
    ZK_CHECK(hipGetLastError());
    k_msm_hist<<<dim3(p.ntiles, p.W), 256, 0, stream>>>(digits, n, p.lo_bits, p.nbins, p.ntiles, hist);
    ZK_CHECK(hipGetLastError());
    ZK_CHECK(prof_mark(PH_SCAN, stream));
    ZK_CHECK(launch_scan(hist, p.nhist, starts, nullptr, scan_sums, stream));
    ZK_CHECK(prof_mark(PH_SCATTER, stream));
    k_msm_scatter_hi<<<dim3(p.ntiles, p.W), 256, 0, stream>>>(digits, n, p.lo_bits, p.nbins, p.ntiles, starts, tmp);
    ZK_CHECK(hipGetLastError());
    k_msm_sort_lo<<<p.W * p.nbins, 256, 0, stream>>>(tmp, starts, p.nbins, p.ntiles, p.lo_bits, p.B, p.W * p.nbins,
                                                     entries, offsets);
    ZK_CHECK(hipGetLastError());
    ZK_CHECK(prof_mark(PH_ACCUMULATE, stream));
    const uint32_t cblocks = (uint32_t)((p.chunks + 255) / 256);
    k_msm_accumulate<<<cblocks, 256, 0, stream>>>(entries, offsets, p.NB, points, buckets, phead, ptail, hb,
                                                  tb, p.chunks);
    ZK_CHECK(hipGetLastError());
    ZK_CHECK(prof_mark(PH_FIXUP, stream));
    k_msm_fixup<<<cblocks, 256, 0, stream>>>(buckets, phead, ptail, hb, tb, p.chunks);
    ZK_CHECK(hipGetLastError());
    ZK_CHECK(prof_mark(PH_REDUCE, stream));
    const uint32_t nseg = p.W * p.T1;
    k_msm_reduce_l1<<<(nseg + 63) / 64, 64, 0, stream>>>(buckets, offsets, p.W, p.B, p.T1, ws1, sum1);
    ZK_CHECK(hipGetLastError());
    k_msm_masked_sums<<<dim3(p.nbits + 1, p.W), 256, 0, stream>>>(ws1, sum1, p.T1, p.nbits, q);
    ZK_CHECK(hipGetLastError());
    ZK_CHECK(prof_mark(PH_FINALIZE, stream));
    k_msm_window_finalize<<<(p.W + 63) / 64, 64, 0, stream>>>(q, p.W, p.nbits, c, slice);
    ZK_CHECK(hipGetLastError());
    ZK_CHECK(prof_mark(PH_NUM, stream));
    return prof_finish();
}

hipError_t msm_collect(uint8_t *const *slices_dev, uint32_t batch, hipStream_t stream) {
    if (!batch) return hipSuccess;
    k_msm_collect<<<(batch + 63) / 64, 64, 0, stream>>>(slices_dev, batch);
    return hipGetLastError();
}

}  // namespace zk

extern "C" {

// cuda/msm.cu:350-432 (caller src/cuda/msm.rs:319-350).  Buffers follow the
// reference contract: res = this MSM's curve slice of
// (windows << window_bits) + 2 * workers Jacobian points; the four index
// buffers hold array_len * windows u32 each and sort_buf 3 * array_len *
// windows u32 (src/cuda/msm.rs:297-311).  This implementation uses those
// buffers as scratch, picks its own signed-digit window (the caller's
// window_bits only sizes the buffers), and leaves the per-window sums in res
// for batch_msm_collect.  It never allocates and never synchronises.
hipError_t msm(void *res, const void *points, const void *scalars, void *indices_buf0, void *indices_buf1,
               void *indices_buf2, void *indices_buf3, void *sort_buf, void *acc_indices_buf, int array_len,
               int windows, int window_bits, int threads, int workers,
               int prepared_sort_indices_temp_storage_bytes, int skip_zero, hipStream_t stream) {
    (void)acc_indices_buf;
    (void)threads;
    (void)prepared_sort_indices_temp_storage_bytes;
    (void)skip_zero;  // zero digits never reach a bucket here; the result is the same
    if (array_len <= 0 || windows <= 0 || window_bits <= 0 || window_bits > 30 || workers <= 0)
        return hipErrorInvalidValue;
    size_t idx_bytes = (size_t)array_len * (size_t)windows * 4;
    uint8_t *regions[5] = {(uint8_t *)sort_buf, (uint8_t *)indices_buf0, (uint8_t *)indices_buf1,
                           (uint8_t *)indices_buf2, (uint8_t *)indices_buf3};
    size_t bytes[5] = {3 * idx_bytes, idx_bytes, idx_bytes, idx_bytes, idx_bytes};
    size_t slice_bytes = (((size_t)windows << window_bits) + 2 * (size_t)workers) * sizeof(G1Jac);
    return msm_windows((uint8_t *)res, slice_bytes, (const G1Aff *)points, (const Fr *)scalars,
                       (uint32_t)array_len, 0, regions, bytes, 5, stream);
}

// cuda/msm.cu:434-467 (caller src/cuda/msm.rs:369-455).  buckets is a device
// array of batch_size device pointers to the curve slices given to msm();
// afterwards each slice starts with that MSM's result (Jacobian, z = 1, or
// the identity (0, 1, 0)).  The remain/next_remain arrays are unused.
hipError_t batch_msm_collect(void *remain_indices, void *remain_acc, void *next_remain_indices,
                             void *next_remain_acc, void *buckets, unsigned workers, unsigned windows,
                             unsigned window_bits, unsigned batch_size, hipStream_t stream) {
    (void)remain_indices;
    (void)remain_acc;
    (void)next_remain_indices;
    (void)next_remain_acc;
    (void)workers;
    (void)windows;
    (void)window_bits;
    if (!buckets && batch_size) return hipErrorInvalidValue;
    return msm_collect((uint8_t *const *)buckets, batch_size, stream);
}

// Phase timing (bench instrumentation): enable / read / reset.  times[k]
// receives the accumulated ms of phase k (count, scan, scatter, accumulate,
// fixup, reduce, window-finalize); returns the number of profiled calls.
void zk_msm_profile_enable(int on) { g_prof.on = on != 0; }
int zk_msm_profile_read(double *times, int ntimes, int reset) {
    for (int k = 0; k < ntimes && k < PH_NUM; k++) times[k] = g_prof.ms[k];
    int calls = (int)g_prof.calls;
    if (reset) {
        for (double &m : g_prof.ms) m = 0;
        g_prof.calls = 0;
    }
    return calls;
}
// The window zk_msm / msm() pick for n points when not told otherwise.
int zk_msm_default_window(int n) { return n > 0 ? (int)default_window((uint32_t)n) : 0; }

// Scratch bytes zk_msm needs for n points at window c (0: default window).
size_t zk_msm_scratch_bytes(int n, int c) {
    if (n <= 0) return 0;
    uint32_t cc = c > 0 ? (uint32_t)c : default_window((uint32_t)n);
    return msm_scratch_bytes((uint32_t)n, cc) + 256;
}

// One MSM with caller-owned scratch (no allocation): the result (Jacobian,
// z = 1, or the identity (0, 1, 0)) lands in out[0..96); out must hold
// ZK_MSM_OUT_BYTES = 96 * 130 bytes (result, header and window sums).
// c = 0 picks the default window for n.
hipError_t zk_msm(void *out, const void *points, const void *scalars, int n, int c, void *scratch,
                  size_t scratch_bytes, hipStream_t stream) {
    if (n <= 0 || !out || !scratch) return hipErrorInvalidValue;
    uint8_t *regions[1] = {(uint8_t *)scratch};
    size_t bytes[1] = {scratch_bytes};
    ZK_CHECK(msm_windows((uint8_t *)out, 96 * 130, (const G1Aff *)points, (const Fr *)scalars, (uint32_t)n,
                         c > 0 ? (uint32_t)c : 0, regions, bytes, 1, stream));
    k_msm_collect_one<<<1, 64, 0, stream>>>((uint8_t *)out);
    return hipGetLastError();
}

}  // extern "C"
