// BN254 G1 multi-scalar multiplication (Pippenger) for gfx950.
//
// Drop-in for `msm` + `batch_msm_collect` (cuda/msm.cu:339-467, host driver
// src/cuda/msm.rs:239-455).  Same buffers and result location: the caller's
// curve slice receives the MSM result as a Jacobian point in its first 96
// bytes after batch_msm_collect.
//
// Pipeline (a batch of MSMs goes through each phase together; buckets are
// numbered (item, window, bucket)):
//   1. digits   one thread per scalar: Montgomery -> canonical, signed c-bit
//               digits d_w in (-2^(c-1), 2^(c-1)], written window-major.
//               Zero digits add nothing.
//   2. sort     two-level counting sort of (bucket, point index | sign):
//               per-tile LDS histograms of the high 8 bucket bits, one global
//               scan, scatter into (item, window, high bits) groups with LDS
//               ranks, then an LDS counting sort of each group on the low 8
//               bits that also emits the bucket offsets.  Writes land in
//               short contiguous runs instead of 4-byte random scatters.
//   3. accumulate  threads own equal chunks of the sorted entries (perfect
//               load balance for any digit distribution); runs of one bucket
//               are summed with XYZZ mixed additions in 29-bit-limb lazy
//               Montgomery arithmetic (ff29.h); runs cut by a chunk border
//               are left as head/tail partials.
//   4. fixup    partials of buckets that span chunks are merged.
//   5. reduce   S_w = sum_b (b+1) B_{w,b}: per segment of L1 buckets a
//               running sum gives (WS, SUM); then
//               S_w = sum_s WS_s + L1 * sum_bit 2^bit * (sum_{s has bit} SUM_s),
//               the masked sums being parallel tree reductions.
//   6. collect  sum_w 2^(c w) S_w by Horner over windows, normalised to z = 1
//               so the bytes are canonical (host cores by default).
// The bucket semantics restate cuda/msm.cu:9-181 (digit d of window w adds
// P_i to bucket (w, d)); signed digits halve the buckets per window.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
#include <thread>
#include <type_traits>
#include <vector>

#include "guard.h"
#include "ff29.h"
#include "host_ec.h"
#include "zkcore.h"

using namespace zk;

namespace {

constexpr uint32_t kL1 = 8;          // buckets per first-level reduction segment
// sorted entries per accumulation thread: 96 (384 B of entries, 128-byte
// aligned) halves the chunk partials the fixup merges against 48, at the
// cost of a slightly longer last wave (2^22 points: fixup 0.31 -> 0.11 ms,
// accumulation +0.05 ms)
constexpr uint32_t kChunk = 96;
// A pass with few sorted entries would leave most of the device idle with
// chunks of kChunk (2^12 points at c = 8: 131k entries, ~1.4k threads, each
// a chain of 96 dependent additions); below kSmallEntries the chunk shrinks
// toward kSmallThreads threads, down to kMinChunk entries.
constexpr uint64_t kSmallEntries = (uint64_t)1 << 21;
constexpr uint64_t kSmallThreads = (uint64_t)1 << 16;
constexpr uint32_t kMinChunk = 8;
constexpr uint32_t kWalk = 32;       // chunk partials one fixup thread sums; longer buckets go to a workgroup
constexpr uint32_t kNone = 0xffffffffu;
constexpr uint32_t kMagic = 0x4d534d32u;  // "MSM2": masked sums present in the slice
// Curve-slice layout left by msm() for batch_msm_collect:
//   [0, 96)          the result (Jacobian), written last
//   [64, 96)         header {magic, c, W, nbits}
//   [128, ...)       Q[w][0..nbits]: W * (nbits + 1) XYZZ masked sums
//   then            S[w]: W Jacobian window sums (device tail only)
constexpr size_t kSliceQ = 128;
// bytes of a slice the layout above can use: W * (nbits + 1) masked sums and
// W window sums stay below 31200 B for every window 2 <= c <= 17 (kL1 >= 4)
constexpr size_t kSliceReserve = 32 * 1024;

// Arrays placed into caller memory, in two lifetimes: phase 1 (the sort) and
// phase 2 (accumulation and reduction).  offsets, entries and the long-bucket
// list live across both;
// the phase-1 temporaries (tmp, hist, starts, scan) are free again in phase 2.
enum MsmArr {
    A_ITEMS, A_SLOTS, A_OFFSETS, A_ENTRIES, A_LONG,                // both phases
    A_TMP, A_HIST, A_STARTS, A_SCAN,                               // phase 1
    A_BUCKETS, A_PHEAD, A_PTAIL, A_HB, A_TB, A_WS1, A_SUM1,        // phase 2
    A_COUNT
};
constexpr int kFirstPhase2 = A_BUCKETS;

constexpr uint32_t kSortTile = 4096;  // scalars per histogram / scatter tile
constexpr uint32_t kPackedMaxN = 1u << 23;  // items up to this size sort with one-word entries

// One MSM of a batch as the kernels see it (built on the device).
struct MsmItemDev {
    const G1Aff *pts;
    const Fr *scalars;
    uint32_t n, tile0, ntiles;  // tile0: the item's first sort tile
    uint64_t dig0;                    // offset of the item's digits (W * scalars before it)
};

// Fixed-base mode: the points of an item are a table T[w * n + i] = 2^(c w) P_i
// (w < Wd = ceil(255 / c) digit windows, built by zk_msm_table_build), so the
// digit of window w of scalar i adds T[w * n + i] to bucket |d| of ONE bucket
// set: sum_i s_i P_i = sum_b b * (sum of the entries with digit b).  The
// additions per scalar stay Wd, but the buckets are reduced once instead of
// per window, which affords a wider window (c = 20: 13 additions per scalar
// against 15 at c = 17).  The sort then runs over n * Wd virtual scalars j =
// w * n + i with up to 2^9 low and 2^10 high bucket bits.
constexpr uint32_t kHistMax = 3840;         // per-tile LDS counters, plain mode (W * nbins, 2 <= c <= 17)
constexpr uint32_t kHistMaxFixed = 13312;   // the same in fixed-base mode (Wd * nbins)
constexpr uint32_t kFixedMinC = 8, kFixedMaxC = 20;
constexpr uint32_t kFixedModelMaxC = 17;  // the widest window zk_msm_table_window picks
constexpr uint32_t kFixedMaxBins = 1024, kFixedMaxKeys = 512;

struct MsmPlan {
    uint32_t count, c, W, B, T1, nbits;
    uint32_t Wd;     // digit windows per scalar (= W, or ceil(255 / c) with W = 1 in fixed-base mode)
    bool fixed;      // fixed-base mode: points are window tables, one bucket set per item
    uint32_t hi_bits, lo_bits, nbins;
    uint64_t ntot, ntiles, nhist, NB, max_entries, chunks;
    uint32_t chunk;  // sorted entries per accumulation thread
    bool packed;  // every item has at most kPackedMaxN scalars: one-word sort entries
    size_t size[A_COUNT];
    uint8_t *ptr[A_COUNT];
};

__host__ __device__ inline uint64_t align256(uint64_t x) { return (x + 255) & ~(uint64_t)255; }

__host__ __device__ inline uint32_t fixed_windows(uint32_t c) { return (255 + c - 1) / c; }

// high bucket bits of fixed-base mode: as many as the per-tile LDS counters
// (Wd * 2^hi <= kHistMaxFixed) and the scatter's bins (<= 2^10) allow, so the
// low-level sort has many groups; the rest (<= 9 bits) are low bits
inline uint32_t fixed_hi_bits(uint32_t c) {
    const uint32_t Wd = fixed_windows(c);
    uint32_t hi = 0;
    while (hi < c - 1 && (2u << hi) <= kFixedMaxBins && (uint64_t)Wd * (2u << hi) <= kHistMaxFixed) hi++;
    return hi;
}

// small: the chunk for few entries (see kSmallEntries), else kChunk
MsmPlan make_plan(const uint32_t *ns, uint32_t count, uint32_t c, bool fixed = false, bool small = true) {
    MsmPlan p;
    memset(&p, 0, sizeof(p));
    p.count = count;
    p.c = c;
    p.fixed = fixed;
    p.Wd = (255 + c - 1) / c;
    p.W = fixed ? 1 : p.Wd;
    p.B = 1u << (c - 1);
    p.T1 = (p.B + kL1 - 1) / kL1;
    p.nbits = log2_exact(p.T1);
    if (fixed) {
        p.hi_bits = fixed_hi_bits(c);
    } else {
        // as many high bits (<= 8) as the per-tile LDS counters allow (W *
        // 2^hi <= kHistMax): the low-level sort gets W * 2^hi groups, so a
        // narrow window still spreads over the device (c = 13 with 16 bins
        // had 320 groups of 16K entries: 2^18 points took 0.62 ms to sort)
        uint32_t hi = 0;
        while (hi < c - 1 && hi < 8 && (uint64_t)p.W * (2u << hi) <= kHistMax) hi++;
        p.hi_bits = hi;
    }
    p.lo_bits = (c - 1) - p.hi_bits;
    p.nbins = 1u << p.hi_bits;
    p.packed = !fixed;  // the virtual indices of fixed-base mode need the two-word entries
    for (uint32_t m = 0; m < count; m++) {
        if (ns[m] > kPackedMaxN) p.packed = false;
        p.ntot += ns[m];
        p.ntiles += (ns[m] + kSortTile - 1) / kSortTile;
    }
    p.NB = (uint64_t)count * p.W * p.B;
    p.nhist = (uint64_t)p.Wd * p.nbins * p.ntiles;
    p.max_entries = p.ntot * p.Wd;
    p.chunk = kChunk;
    if (small && p.max_entries < kSmallEntries) {
        const uint64_t ch = (p.max_entries + kSmallThreads - 1) / kSmallThreads;
        p.chunk = ch < kMinChunk ? kMinChunk : (uint32_t)ch;
    }
    p.chunks = (p.max_entries + p.chunk - 1) / p.chunk;
    p.size[A_ITEMS] = (size_t)(count + 1) * sizeof(MsmItemDev);
    p.size[A_SLOTS] = (size_t)count * sizeof(uint8_t *);
    p.size[A_OFFSETS] = (size_t)(p.NB + 1) * 4;
    // digits (window-major per item, ntot * W u32) share this array with the entries
    p.size[A_ENTRIES] = (size_t)p.max_entries * 4;
    p.size[A_TMP] = (size_t)p.max_entries * (p.packed ? 4 : 8);
    p.size[A_HIST] = (size_t)p.nhist * 4;
    p.size[A_STARTS] = (size_t)(p.nhist + 1) * 4;
    p.size[A_SCAN] = ((size_t)(p.nhist + 1 + 1023) / 1024 + 1) * 4;
    p.size[A_BUCKETS] = (size_t)p.NB * sizeof(X29S);
    p.size[A_PHEAD] = (size_t)p.chunks * sizeof(X29S);
    p.size[A_PTAIL] = (size_t)p.chunks * sizeof(X29S);
    p.size[A_HB] = (size_t)p.chunks * 4;
    p.size[A_TB] = (size_t)p.chunks * 4;
    p.size[A_WS1] = (size_t)count * p.W * p.T1 * sizeof(X29S);
    p.size[A_SUM1] = (size_t)count * p.W * p.T1 * sizeof(X29S);
    // [count, (bucket, first chunk, last chunk) per bucket spanning > kWalk chunks]
    p.size[A_LONG] = (size_t)(1 + 3 * (p.chunks / kWalk + 1)) * 4;
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
    if (!place_range(p, A_ITEMS, kFirstPhase2, segs, nseg)) return false;
    for (int a = A_TMP; a < kFirstPhase2; a++) {  // phase-1 temporaries become free
        segs[nseg].ptr = p.ptr[a];
        segs[nseg].left = align256(p.size[a]);
        nseg++;
    }
    return place_range(p, kFirstPhase2, A_COUNT, segs, nseg);
}

// Tuned windows per size class ceil(log2 n); 0 = not tuned.  Written by the
// autotuner (zk_msm_autotune / zk_msm_set_window), read by every MSM: atomic,
// since those calls may come from other threads than the MSMs (the window
// changes only the speed of an MSM, never its result).
std::atomic<uint8_t> g_window[32];

// Window for a single MSM of n points: the tuned entry for its size class,
// else the model log2(n) - 5 clamped to [6, 17] (bucket reduction ~ 2^c per
// window against n additions per window), with c = 15 for 2^18..2^20 points:
// there c = 13 and 14 leave a top window of 7 and 2 bits, whose few buckets
// take the n / 2 or more entries of that window (measured at 2^18, per MSM:
// c = 13 1.99 ms, c = 15 1.29 ms).
uint32_t default_window(uint32_t n) {
    uint32_t lg = log2_exact(n ? n : 1);
    const uint8_t tuned = lg < 32 ? g_window[lg].load(std::memory_order_relaxed) : 0;
    if (tuned) return tuned;
    // measured per size (zk_msm, one box): 15 for 2^18..2^19 (13 / 14 leave a
    // narrow top window whose few buckets take half its entries), 16 for
    // 2^20 (2.06 against 2.11 ms at 15), 17 from 2^21 (3.22 against 3.66 ms
    // at 16), else lg - 5
    int c = lg >= 21 ? 17 : (lg == 20 ? 16 : (lg >= 18 ? 15 : (int)lg - 5));
    if (c < 6) c = 6;
    if (c > 17) c = 17;
    return (uint32_t)c;
}

__device__ __forceinline__ uint32_t scalar_bits(const uint32_t s[8], uint32_t start, uint32_t c) {
    const uint32_t limb = start >> 5, off = start & 31;
    // the two limbs by selects: a dynamic index would put s in scratch memory
    uint32_t lo = 0, hi = 0;
#pragma unroll
    for (uint32_t k = 0; k < 8; k++) {
        lo = k == limb ? s[k] : lo;
        hi = k == limb + 1 ? s[k] : hi;
    }
    uint32_t v = lo >> off;
    if (off) v |= hi << (32 - off);
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

// item of a global sort-tile index: the last item whose first tile is <= x
__device__ __forceinline__ uint32_t item_of(const MsmItemDev *__restrict__ items, uint32_t count, uint32_t x) {
    uint32_t lo = 0, hi = count;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (items[mid].tile0 <= x) lo = mid; else hi = mid;
    }
    return lo;
}

// The item table (with the prefixes computed on the host) travels in kernel
// arguments, kItemsPerLaunch items per launch, so no host buffer is copied.
constexpr uint32_t kItemsPerLaunch = 16;
struct ItemPack {
    MsmItemDev it[kItemsPerLaunch];
    uint8_t *slot[kItemsPerLaunch];  // the item's curve slice
    uint32_t first, k;
};
__global__ void k_msm_put_items(ItemPack pk, MsmItemDev *__restrict__ items, uint8_t **__restrict__ slots,
                                uint32_t *__restrict__ long_count) {
    uint32_t j = threadIdx.x;
    if (pk.first == 0 && j == 0) *long_count = 0;
    if (j >= pk.k) return;
    items[pk.first + j] = pk.it[j];
    slots[pk.first + j] = pk.slot[j];
}

// hist index of (item m, digit window w, bin, tile-in-item).  Plain mode:
// groups (m, w, bin) of ntiles counters each; fixed-base mode (one bucket
// set per item): groups (m, bin) of Wd * ntiles counters, ordered (w, tile)
// inside, so the virtual scalars w * n + i stay in order.  Groups are
// contiguous and ordered, so their exclusive scan gives every group's start.
__device__ __forceinline__ uint64_t hist_index(const MsmItemDev &it, uint32_t Wd, uint32_t nbins, bool fixed,
                                               uint32_t w, uint32_t bin, uint32_t tl) {
    const uint64_t h0 = (uint64_t)Wd * nbins * it.tile0;
    return fixed ? h0 + ((uint64_t)bin * Wd + w) * it.ntiles + tl : h0 + ((uint64_t)w * nbins + bin) * it.ntiles + tl;
}
// counters per sort group
__device__ __forceinline__ uint32_t group_tiles(const MsmItemDev &it, uint32_t Wd, bool fixed) {
    return fixed ? Wd * it.ntiles : it.ntiles;
}

// Digits and their histogram in one pass, one block per sort tile (4096
// scalars of one item, 4 per thread, coalesced).  digits[dig0 + w * n + i] =
// 0 for a zero digit, else |d| | (d < 0) << 31 with |d| in [1, 2^(c-1)]
// (bucket |d| - 1); per digit window the counts of the high bucket bits are
// kept in LDS (Wd * nbins <= HMAX counters: 3840 for every plain window 2 <=
// c <= 17, 13312 in fixed-base mode) and written once per tile, so the digits
// are not read back for the histogram.
constexpr uint32_t kDigitThreads = 1024;
template <uint32_t HMAX>
__global__ __launch_bounds__(kDigitThreads) void k_msm_digits_hist(const MsmItemDev *__restrict__ items, uint32_t count,
                                                         uint32_t c, uint32_t Wd, uint32_t lo_bits, uint32_t nbins,
                                                         uint32_t fixed, uint32_t *__restrict__ digits,
                                                         uint32_t *__restrict__ hist) {
    __shared__ uint32_t h[HMAX];
    const uint32_t m = item_of(items, count, blockIdx.x);
    const MsmItemDev it = items[m];
    const uint32_t tl = blockIdx.x - it.tile0;
    const uint32_t nh = Wd * nbins;
    for (uint32_t k = threadIdx.x; k < nh; k += blockDim.x) h[k] = 0;
    __syncthreads();
    uint32_t *dst = digits + it.dig0;
    const uint32_t end = (tl + 1) * kSortTile < it.n ? (tl + 1) * kSortTile : it.n;
    for (uint32_t i = tl * kSortTile + threadIdx.x; i < end; i += blockDim.x) {
        const Fr s = fe_from_mont<FrParams>(fe_load<FrParams>(&it.scalars[i]));
        uint32_t carry = 0;
        for (uint32_t w = 0; w < Wd; w++) {
            const int32_t d = signed_digit(s.v, w, c, carry);
            const uint32_t a = (uint32_t)(d < 0 ? -d : d);
            dst[(size_t)w * it.n + i] = d ? (a | (d < 0 ? 0x80000000u : 0u)) : 0u;
            if (d) atomicAdd(&h[w * nbins + ((a - 1) >> lo_bits)], 1u);
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nh; k += blockDim.x)
        hist[hist_index(it, Wd, nbins, fixed != 0, k / nbins, k % nbins, tl)] = h[k];
}

// exclusive scan of x[0..256) in LDS by 256 threads (4 waves): a
// shuffle scan inside each wave, the 4 wave totals through t[0..4), two
// barriers (x[k] = 0 past the used range)
__device__ __forceinline__ void lds_exclusive_scan256(uint32_t *x, uint32_t *t) {
    const uint32_t i = threadIdx.x, lane = i & 63, wv = i >> 6;
    const uint32_t v = x[i];
    uint32_t s = v;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t o = (uint32_t)__shfl_up((int)s, d, 64);
        if (lane >= d) s += o;
    }
    if (lane == 63) t[wv] = s;
    __syncthreads();
    uint32_t off = 0;
    for (uint32_t k = 0; k < wv; k++) off += t[k];
    x[i] = s - v + off;
    __syncthreads();
}

// exclusive scan of x[0..K) (K = 256 R) by 256 threads: thread i owns the R
// consecutive counters from i R, the 256 per-thread totals are scanned as
// above
template <uint32_t K>
__device__ __forceinline__ void lds_exclusive_scan(uint32_t *x, uint32_t *t) {
    constexpr uint32_t R = K / 256;
    static_assert(R >= 1 && R * 256 == K, "K must be a multiple of 256");
    if constexpr (R == 1) {
        lds_exclusive_scan256(x, t);
    } else {
        const uint32_t i = threadIdx.x, lane = i & 63, wv = i >> 6;
        uint32_t loc[R];
        uint32_t v = 0;
#pragma unroll
        for (uint32_t r = 0; r < R; r++) {
            loc[r] = x[i * R + r];
            v += loc[r];
        }
        uint32_t s = v;
#pragma unroll
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t o = (uint32_t)__shfl_up((int)s, d, 64);
            if (lane >= d) s += o;
        }
        if (lane == 63) t[wv] = s;
        __syncthreads();
        uint32_t run = s - v;
        for (uint32_t k = 0; k < wv; k++) run += t[k];
#pragma unroll
        for (uint32_t r = 0; r < R; r++) {
            x[i * R + r] = run;
            run += loc[r];
        }
        __syncthreads();
    }
}

// Entries between the two sort levels: (low bucket bits, scalar index |
// sign << 31), as a pair of words, or packed into one word when every item
// has at most 2^23 scalars (sign | low bits << 23 | index), which halves the
// traffic of the pass between the two levels.
template <bool kPacked> struct TmpEnt;
template <> struct TmpEnt<false> {
    using T = uint2;
    static __device__ __forceinline__ T make(uint32_t lo, uint32_t v) { return make_uint2(lo, v); }
    static __device__ __forceinline__ uint32_t key(const T &e) { return e.x; }
    static __device__ __forceinline__ uint32_t val(const T &e) { return e.y; }
};
template <> struct TmpEnt<true> {
    using T = uint32_t;
    static __device__ __forceinline__ T make(uint32_t lo, uint32_t v) {
        return (v & 0x80000000u) | (lo << 23) | (v & 0x7fffffu);
    }
    static __device__ __forceinline__ uint32_t key(const T &e) { return (e >> 23) & 0xffu; }
    static __device__ __forceinline__ uint32_t val(const T &e) { return e & 0x807fffffu; }
};
// per (tile, digit window): place each nonzero digit into its sort group
// ((item, window, bin), or (item, bin) in fixed-base mode); tmp entry = (low
// bucket bits, index | sign), the index being the scalar's (plain) or the
// virtual scalar's w * n + i (fixed-base: the table row of the point).  The
// tile is first counting-sorted by bin in LDS (the bin counts are the
// histogram's), then written out bin run by bin run, so consecutive lanes
// store to consecutive addresses instead of 256 scattered writes per wave.
// NBMAX: bins the LDS arrays hold (256, or 1024 in fixed-base mode).
template <bool kPacked, uint32_t NBMAX>
__global__ __launch_bounds__(256) void k_msm_scatter_hi(const MsmItemDev *__restrict__ items, uint32_t count,
                                                        uint32_t Wd, const uint32_t *__restrict__ digits,
                                                        uint32_t lo_bits, uint32_t nbins, uint32_t fixed,
                                                        const uint32_t *__restrict__ hist,
                                                        const uint32_t *__restrict__ starts,
                                                        typename TmpEnt<kPacked>::T *__restrict__ tmp) {
    using E = TmpEnt<kPacked>;
    using BinT = typename std::conditional<(NBMAX > 256), uint16_t, uint8_t>::type;
    __shared__ uint32_t gstart[NBMAX], loff[NBMAX], lcur[NBMAX], sc[4];
    __shared__ typename E::T stage[kSortTile];
    __shared__ BinT sbin[kSortTile];
    const uint32_t w = blockIdx.y;
    const uint32_t m = item_of(items, count, blockIdx.x);
    const MsmItemDev it = items[m];
    const uint32_t tl = blockIdx.x - it.tile0;
    const uint32_t tid = threadIdx.x;
    for (uint32_t k = tid; k < NBMAX; k += 256) {
        uint32_t cnt = 0;
        if (k < nbins) {
            const uint64_t hi = hist_index(it, Wd, nbins, fixed != 0, w, k, tl);
            gstart[k] = starts[hi];
            cnt = hist[hi];
        }
        loff[k] = cnt;  // bins past nbins count 0
    }
    __syncthreads();
    lds_exclusive_scan<NBMAX>(loff, sc);
    for (uint32_t k = tid; k < NBMAX; k += 256) lcur[k] = loff[k];
    __syncthreads();
    const uint32_t *dw = digits + it.dig0 + (size_t)w * it.n;
    const uint32_t vbase = fixed ? w * it.n : 0u;  // fixed-base: the table row of (w, i) is w * n + i
    const uint32_t lo_mask = (1u << lo_bits) - 1;
    const uint32_t begin = tl * kSortTile;
    const uint32_t end = begin + kSortTile < it.n ? begin + kSortTile : it.n;
    for (uint32_t i = begin + tid; i < end; i += blockDim.x) {
        uint32_t d = dw[i];
        if (d) {
            uint32_t b = (d & 0x7fffffffu) - 1;
            uint32_t p = atomicAdd(&lcur[b >> lo_bits], 1u);
            stage[p] = E::make(b & lo_mask, (vbase + i) | (d & 0x80000000u));
            sbin[p] = (BinT)(b >> lo_bits);
        }
    }
    __syncthreads();
    const uint32_t tot = lcur[nbins - 1];  // the last bin's cursor ends at the tile's total
    for (uint32_t p = tid; p < tot; p += blockDim.x) {
        const uint32_t k = sbin[p];
        tmp[gstart[k] + (p - loff[k])] = stage[p];
    }
}

// per sort group ((item, window, bin), or (item, bin) in fixed-base mode),
// one block of 1024 threads: counting sort on the low bucket bits, and the
// bucket offsets.  The sorted group fills exactly its range [gs, ge) of the
// entries.  A group of at most kGroupMax entries (a 2^22-point MSM's groups
// hold ~16k, its top window's ~22k) is sorted whole: loaded once into
// registers (24 entries per thread, coalesced), its key counts taken by LDS
// atomics and scanned into the bucket offsets, each entry placed in LDS at its
// key's cursor, and the group written out in order, fully coalesced.  A
// larger group takes a counting pass over the whole group for the bucket
// offsets, then tiles of kGroupMax entries, each counting-sorted in LDS (its
// keys kept beside it) and written out bucket run by bucket run.  KEYS:
// low-bit values the LDS arrays hold (256, or 512 in fixed-base mode).
constexpr uint32_t kGroupThreads = 1024;
constexpr uint32_t kGroupPer = 24;  // entries per thread
constexpr uint32_t kGroupMax = kGroupThreads * kGroupPer;

// exclusive scan of x[0..K) in LDS (K <= 1024) by a block of 1024 threads: a
// shuffle scan inside each wave, the 16 wave totals through t[0..16)
template <uint32_t K>
__device__ __forceinline__ void lds_exclusive_scan1024(uint32_t *x, uint32_t *t) {
    static_assert(K <= kGroupThreads, "one counter per thread");
    const uint32_t i = threadIdx.x, lane = i & 63, wv = i >> 6;
    const uint32_t v = i < K ? x[i] : 0u;
    uint32_t s = v;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t o = (uint32_t)__shfl_up((int)s, d, 64);
        if (lane >= d) s += o;
    }
    if (lane == 63) t[wv] = s;
    __syncthreads();
    uint32_t off = 0;
    for (uint32_t k = 0; k < wv; k++) off += t[k];
    if (i < K) x[i] = s - v + off;
    __syncthreads();
}

template <bool kPacked, uint32_t KEYS>
__global__ __launch_bounds__(kGroupThreads) void k_msm_sort_group(
    const MsmItemDev *__restrict__ items, uint32_t W, uint32_t Wd, uint32_t nbins, uint32_t fixed,
    const typename TmpEnt<kPacked>::T *__restrict__ tmp, const uint32_t *__restrict__ starts, uint32_t lo_bits,
    uint32_t B, uint32_t ngroups, uint32_t *__restrict__ entries, uint32_t *__restrict__ offsets) {
    using E = TmpEnt<kPacked>;
    using KeyT = typename std::conditional<(KEYS > 256), uint16_t, uint8_t>::type;
    __shared__ uint32_t h[KEYS], cur[KEYS], th[KEYS], tcur[KEYS], sc[kGroupThreads / 64];
    __shared__ uint32_t stage[kGroupMax];
    __shared__ KeyT skey[kGroupMax];
    const uint32_t g = blockIdx.x;
    const uint32_t m = g / (W * nbins), r = g % (W * nbins);
    const uint32_t w = r / nbins, bin = r % nbins;
    const MsmItemDev it = items[m];
    const uint64_t idx = hist_index(it, Wd, nbins, fixed != 0, w, bin, 0);
    const uint32_t gs = starts[idx];
    const uint32_t ge = starts[idx + group_tiles(it, Wd, fixed != 0)];  // the next group's start (or the total)
    const uint32_t tid = threadIdx.x;
    const uint32_t nlo = 1u << lo_bits;
    const uint64_t base = ((uint64_t)m * W + w) * B + ((uint64_t)bin << lo_bits);
    if (g == ngroups - 1 && tid == 0) offsets[base + nlo] = ge;  // = total: offsets[NB]
    typename E::T v[kGroupPer];
    if (ge - gs <= kGroupMax) {
        const uint32_t cnt = ge - gs;
        for (uint32_t k = tid; k < KEYS; k += kGroupThreads) h[k] = 0;
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < kGroupPer; k++) {
            const uint32_t q = k * kGroupThreads + tid;
            if (q < cnt) {
                v[k] = tmp[gs + q];
                atomicAdd(&h[E::key(v[k])], 1u);
            }
        }
        __syncthreads();
        lds_exclusive_scan1024<KEYS>(h, sc);
        for (uint32_t k = tid; k < nlo; k += kGroupThreads) offsets[base + k] = gs + h[k];
        __syncthreads();  // h becomes the cursors
#pragma unroll
        for (uint32_t k = 0; k < kGroupPer; k++) {
            const uint32_t q = k * kGroupThreads + tid;
            if (q < cnt) stage[atomicAdd(&h[E::key(v[k])], 1u)] = E::val(v[k]);
        }
        __syncthreads();
        for (uint32_t q = tid; q < cnt; q += kGroupThreads) entries[gs + q] = stage[q];
        return;
    }
    // a larger group: bucket offsets from a counting pass over the group
    for (uint32_t k = tid; k < KEYS; k += kGroupThreads) cur[k] = 0;
    __syncthreads();
    for (uint32_t p = gs + tid; p < ge; p += kGroupThreads) atomicAdd(&cur[E::key(tmp[p])], 1u);
    __syncthreads();
    lds_exclusive_scan1024<KEYS>(cur, sc);
    for (uint32_t k = tid; k < KEYS; k += kGroupThreads) cur[k] += gs;
    __syncthreads();
    for (uint32_t k = tid; k < nlo; k += kGroupThreads) offsets[base + k] = cur[k];
    for (uint32_t t0 = gs; t0 < ge; t0 += kGroupMax) {
        const uint32_t cnt = ge - t0 < kGroupMax ? ge - t0 : kGroupMax;
        for (uint32_t k = tid; k < KEYS; k += kGroupThreads) th[k] = 0;
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < kGroupPer; k++) {
            const uint32_t q = k * kGroupThreads + tid;
            if (q < cnt) {
                v[k] = tmp[t0 + q];
                atomicAdd(&th[E::key(v[k])], 1u);
            }
        }
        __syncthreads();
        for (uint32_t k = tid; k < KEYS; k += kGroupThreads) h[k] = th[k];
        __syncthreads();
        lds_exclusive_scan1024<KEYS>(h, sc);  // h: the tile's key offsets
        for (uint32_t k = tid; k < KEYS; k += kGroupThreads) tcur[k] = h[k];
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < kGroupPer; k++) {
            const uint32_t q = k * kGroupThreads + tid;
            if (q < cnt) {
                const uint32_t key = E::key(v[k]);
                const uint32_t at = atomicAdd(&tcur[key], 1u);
                stage[at] = E::val(v[k]);
                skey[at] = (KeyT)key;
            }
        }
        __syncthreads();
        for (uint32_t q = tid; q < cnt; q += kGroupThreads) {
            const uint32_t key = skey[q];
            entries[cur[key] + (q - h[key])] = stage[q];
        }
        __syncthreads();
        for (uint32_t k = tid; k < KEYS; k += kGroupThreads) cur[k] += th[k];
        __syncthreads();
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

// Buckets are numbered (item, window, bucket) = (m * W + w) * B + b over the
// whole batch; the entries of item m occupy [off[m * W * B], off[(m + 1) * W * B]).
__global__ __launch_bounds__(256) void k_msm_accumulate(const uint32_t *__restrict__ entries,
                                                        const uint32_t *__restrict__ off, uint32_t NB, uint32_t WB,
                                                        const MsmItemDev *__restrict__ items,
                                                        X29S *__restrict__ buckets, X29S *__restrict__ phead,
                                                        X29S *__restrict__ ptail, uint32_t *__restrict__ hb,
                                                        uint32_t *__restrict__ tb, uint64_t nchunks, uint32_t chunk) {
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nchunks) return;
    const uint32_t total = off[NB];
    const uint32_t s = (uint32_t)(t * chunk);
    uint32_t head_bucket = kNone, tail_bucket = kNone;
    if (s < total) {
        const uint32_t e = s + chunk < total ? s + chunk : total;
        uint32_t b = find_bucket(off, NB, s);
        uint32_t bend = off[b + 1];
        bool first = true, before = off[b] < s;
        // the bucket sum in 29-bit limbs (see ff29.h), converted at each store
        X29 acc;
        bool empty;
        acc.x = f29_one();
        acc.y = acc.x;
        x29_clear(acc, empty);
        // item (and point array) of the entry being prefetched
        uint32_t m = b / WB;
        uint32_t mend = off[(m + 1) * WB];
        const G1Aff *pts = items[m].pts;
        // one uniform loop over the chunk (every lane runs e - s steps); the
        // next entry and its point are loaded one step ahead
        uint32_t ecur = entries[s];
        G1Aff qcur = aff_load(&pts[ecur & 0x7fffffffu]);
        for (uint32_t pos = s; pos < e; pos++) {
            const bool more = pos + 1 < e;
            if (more && pos + 1 >= mend) {
                // the next entry belongs to a later item (pos + 1 < total
                // bounds m + 1 by the item count)
                do {
                    m++;
                    mend = off[(m + 1) * WB];
                } while (mend <= pos + 1);
                pts = items[m].pts;
            }
            const uint32_t enext = more ? entries[pos + 1] : 0u;
            const G1Aff qnext = more ? aff_load(&pts[enext & 0x7fffffffu]) : qcur;
            if (!aff_is_id(qcur)) {
                Fp y = (ecur >> 31) ? fe_neg<FpParams>(qcur.y) : qcur.y;
                x29_acc_aff(acc, empty, f29_from32(qcur.x), f29_from32(y));
            }
            if (pos + 1 == bend || !more) {
                // the run of bucket b ends at this step (stored in the 29-bit
                // form: the fixup and the reduction continue in it)
                if (first && before) {
                    x29_store_acc(&phead[t], acc, empty);
                    head_bucket = b;
                } else if (bend > e) {
                    x29_store_acc(&ptail[t], acc, empty);
                    tail_bucket = b;
                } else {
                    x29_store_acc(&buckets[b], acc, empty);
                }
                first = false;
                x29_clear(acc, empty);
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

// merges the partial sums of buckets that span several chunks.  A bucket b
// whose run starts in chunk t (tail partial) ends in chunk t' = (off[b+1]-1) /
// chunk, and chunks t+1..t' all hold head partials of b.  Spans up to kWalk
// chunks are summed by the thread; longer ones (few: their count is at most
// nchunks / kWalk) are listed for k_msm_fixup_long, one workgroup each, so a
// heavy bucket (e.g. the multiplicities of a lookup, all in a handful of
// buckets) never becomes a serial chain.
__global__ __launch_bounds__(256) void k_msm_fixup(X29S *__restrict__ buckets, const X29S *__restrict__ phead,
                                                   const X29S *__restrict__ ptail, const uint32_t *__restrict__ tb,
                                                   const uint32_t *__restrict__ off, uint32_t *__restrict__ longs,
                                                   uint64_t nchunks, uint32_t chunk) {
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nchunks) return;
    uint32_t b = tb[t];
    if (b == kNone) return;
    const uint64_t last = (uint64_t)(off[b + 1] - 1) / chunk;
    if (last - t > kWalk) {
        uint32_t i = atomicAdd(&longs[0], 1u);
        longs[1 + 3 * i] = b;
        longs[2 + 3 * i] = (uint32_t)t;
        longs[3 + 3 * i] = (uint32_t)last;
        return;
    }
    bool ae, qe;
    X29 acc = x29_load(&ptail[t], ae);
    for (uint64_t u = t + 1; u <= last; u++) {
        const X29 q = x29_load(&phead[u], qe);
        x29_add(acc, ae, q, qe);
    }
    x29_store(&buckets[b], acc, ae);
}

// one workgroup per listed long bucket: ptail[t] + sum of phead[t+1..last]
__global__ __launch_bounds__(256) void k_msm_fixup_long(X29S *__restrict__ buckets,
                                                        const X29S *__restrict__ phead,
                                                        const X29S *__restrict__ ptail,
                                                        const uint32_t *__restrict__ longs) {
    __shared__ X29S sh[256];
    const uint32_t count = longs[0];
    for (uint32_t i = blockIdx.x; i < count; i += gridDim.x) {
        const uint32_t b = longs[1 + 3 * i], t = longs[2 + 3 * i], last = longs[3 + 3 * i];
        bool ae = true, qe;
        X29 acc;
        if (threadIdx.x == 0) acc = x29_load(&ptail[t], ae);
        for (uint32_t u = t + 1 + threadIdx.x; u <= last; u += blockDim.x) {
            const X29 q = x29_load(&phead[u], qe);
            x29_add(acc, ae, q, qe);
        }
        x29_store(&sh[threadIdx.x], acc, ae);
        __syncthreads();
        for (uint32_t st = 128; st; st >>= 1) {
            if (threadIdx.x < st) {
                X29 l = x29_load(&sh[threadIdx.x], ae);
                const X29 r = x29_load(&sh[threadIdx.x + st], qe);
                x29_add(l, ae, r, qe);
                x29_store(&sh[threadIdx.x], l, ae);
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            const X29 r = x29_load(&sh[0], ae);
            x29_store(&buckets[b], r, ae);
        }
        __syncthreads();
    }
}

// buckets that some chunk starts inside and no chunk owns fully are
// covered above; a bucket whose whole range lies in one chunk was written
// directly.  Empty buckets are never read (their count is zero).

// ---- reduction -------------------------------------------------------------
// one thread per (item window, segment of kL1 buckets); W = items * windows
__global__ __launch_bounds__(64) void k_msm_reduce_l1(const X29S *__restrict__ buckets, const uint32_t *__restrict__ off,
                                uint32_t W, uint32_t B, uint32_t T1, X29S *__restrict__ ws1,
                                X29S *__restrict__ sum1) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= W * T1) return;
    uint32_t w = g / T1, s = g % T1;
    X29 run, acc;
    bool re = true, ae = true, be;
    uint32_t lo = s * kL1;
    uint32_t hi = lo + kL1 < B ? lo + kL1 : B;
    for (uint32_t b = hi; b-- > lo;) {
        uint32_t idx = w * B + b;
        if (off[idx + 1] != off[idx]) {
            const X29 q = x29_load(&buckets[idx], be);
            x29_add(run, re, q, be);
        }
        x29_add(acc, ae, run, re);
    }
    x29_store(&ws1[g], acc, ae);
    x29_store(&sum1[g], run, re);
}

// the item window g's masked sum Q[bit] in its curve slice
__device__ __forceinline__ G1Xyzz *slot_q(uint8_t *const *slots, uint32_t g, uint32_t W, uint32_t nbits,
                                          uint32_t bit) {
    G1Xyzz *q = reinterpret_cast<G1Xyzz *>(slots[g / W] + kSliceQ);
    return &q[(size_t)(g % W) * (nbits + 1) + bit];
}

// an X29 point and its identity flag from lane ^ m
__device__ __forceinline__ void x29_shfl_xor(X29 &o, bool &oe, const X29 &p, bool pe, int m) {
#pragma unroll
    for (int i = 0; i < 9; i++) {
        o.x.l[i] = (uint32_t)__shfl_xor((int)p.x.l[i], m);
        o.y.l[i] = (uint32_t)__shfl_xor((int)p.y.l[i], m);
        o.zz.l[i] = (uint32_t)__shfl_xor((int)p.zz.l[i], m);
        o.zzz.l[i] = (uint32_t)__shfl_xor((int)p.zzz.l[i], m);
    }
    oe = __shfl_xor((int)pe, m) != 0;
}
__device__ __forceinline__ X29 x29_select(bool c, const X29 &a, const X29 &b) {
    X29 r;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        r.x.l[i] = c ? a.x.l[i] : b.x.l[i];
        r.y.l[i] = c ? a.y.l[i] : b.y.l[i];
        r.zz.l[i] = c ? a.zz.l[i] : b.zz.l[i];
        r.zzz.l[i] = c ? a.zzz.l[i] : b.zzz.l[i];
    }
    return r;
}

// Level-1 reduction with the masked sums folded in (T1 >= 64): one wave per 64
// consecutive segments s = 64 h + lane of an item window g.  Each lane forms
// its segment's run (plain sum of its kL1 buckets) and acc (weighted sum) as
// k_msm_reduce_l1 does; then six butterfly steps over the lanes: at step k a
// lane with bit k clear adds its partner's run, a lane with bit k set its
// partner's acc, so every lane does one point addition per step.  Afterwards
// lane 0 holds the wave's total run T, lane 2^j (j < 6) the sum of the runs
// of the lanes with bit j set, and lane 63 the total acc A.  They go to
// part[(g * T1/64 + h) * 8 + {0, 1 + j, 7}] for k_msm_fold_sums.
__global__ __launch_bounds__(64) void k_msm_reduce_fold(const X29S *__restrict__ buckets,
                                                        const uint32_t *__restrict__ off, uint32_t B, uint32_t T1,
                                                        X29S *__restrict__ part) {
    const uint32_t lane = threadIdx.x, hw = T1 >> 6;
    const uint32_t g = blockIdx.x / hw, h = blockIdx.x % hw;
    const uint32_t s = h * 64 + lane;
    X29 run, acc;
    bool re = true, ae = true, be;
    const uint32_t lo = s * kL1;
    const uint32_t hi = lo + kL1 < B ? lo + kL1 : B;
    for (uint32_t b = hi; b-- > lo;) {
        const uint32_t idx = g * B + b;
        if (off[idx + 1] != off[idx]) {
            const X29 q = x29_load(&buckets[idx], be);
            x29_add(run, re, q, be);
        }
        x29_add(acc, ae, run, re);
    }
#pragma unroll 1
    for (int k = 0; k < 6; k++) {
        const bool up = (lane >> k) & 1u;
        // the partner (bit k flipped) needs the quantity this lane does not update
        X29 o;
        bool oe;
        x29_shfl_xor(o, oe, x29_select(up, run, acc), up ? re : ae, 1 << k);
        X29 a = x29_select(up, acc, run);
        bool e = up ? ae : re;
        x29_add(a, e, o, oe);
        if (up) {
            acc = a;
            ae = e;
        } else {
            run = a;
            re = e;
        }
    }
    X29S *dst = part + ((size_t)g * hw + h) * 8;
    if (lane == 63)
        x29_store(&dst[7], acc, ae);
    else if (lane == 0)
        x29_store(&dst[0], run, re);
    else if ((lane & (lane - 1)) == 0)
        x29_store(&dst[1 + __builtin_ctz(lane)], run, re);
}

// grid (nbits + 1, gw), one wave each: the item window's masked sum Q[o] from
// the partials of k_msm_reduce_fold (hw = T1 / 64 per window).  o < 6: sum of
// the waves' lane-bit partials j = o; 6 <= o < nbits: sum of the totals T_h
// of the waves with bit o - 6 of h set; o = nbits: sum of the waves' A.
__global__ __launch_bounds__(64) void k_msm_fold_sums(const X29S *__restrict__ part, uint32_t T1, uint32_t nbits,
                                                      uint32_t W, uint8_t *const *__restrict__ slots) {
    const uint32_t o = blockIdx.x, g = blockIdx.y, lane = threadIdx.x, hw = T1 >> 6;
    const uint32_t slot = o < 6 ? 1 + o : (o < nbits ? 0 : 7);
    X29 acc;
    bool ae = true, qe;
    for (uint32_t h = lane; h < hw; h += 64) {
        if (o < 6 || o == nbits || ((h >> (o - 6)) & 1u)) {
            const X29 q = x29_load(&part[((size_t)g * hw + h) * 8 + slot], qe);
            x29_add(acc, ae, q, qe);
        }
    }
#pragma unroll 1
    for (int k = 0; k < 6; k++) {
        X29 q;
        x29_shfl_xor(q, qe, acc, ae, 1 << k);
        x29_add(acc, ae, q, qe);
    }
    if (lane == 0) xyzz_store(slot_q(slots, g, W, nbits, o), x29_to_xyzz(acc, ae));
}

// grid (nbits + 1, items * W, parts): q < nbits -> sum of sum1[g][s] over s
// with bit q set; q == nbits -> sum of ws1[g][s] over all s.  With one part
// the sum goes into the item's curve slice at Q[w][q]; with several, block z
// sums the z-th slice of s into qpart[(g * (nbits + 1) + q) * parts + z] and
// k_msm_masked_merge adds the parts (a single MSM has only (nbits+1) * W
// sums, too few workgroups to fill the device).
// LDS tree over the 256 threads' X29 values in sh; the sum ends in sh[0]
__device__ __forceinline__ void x29_block_tree(X29S *sh) {
    bool ae, qe;
    for (uint32_t st = 128; st; st >>= 1) {
        if (threadIdx.x < st) {
            X29 l = x29_load(&sh[threadIdx.x], ae);
            const X29 r = x29_load(&sh[threadIdx.x + st], qe);
            x29_add(l, ae, r, qe);
            x29_store(&sh[threadIdx.x], l, ae);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_msm_masked_sums(const X29S *__restrict__ ws1,
                                                         const X29S *__restrict__ sum1, uint32_t T1,
                                                         uint32_t nbits, uint32_t W,
                                                         uint8_t *const *__restrict__ slots,
                                                         X29S *__restrict__ qpart) {
    __shared__ X29S sh[256];
    const uint32_t bit = blockIdx.x, g = blockIdx.y, parts = gridDim.z, part = blockIdx.z;
    const uint32_t span = T1 / parts, s0 = part * span;
    const X29S *src = (bit == nbits ? ws1 : sum1) + (size_t)g * T1;
    X29 acc;
    bool ae = true, qe;
    for (uint32_t s = s0 + threadIdx.x; s < s0 + span; s += blockDim.x) {
        if (bit == nbits || ((s >> bit) & 1u)) {
            const X29 q = x29_load(&src[s], qe);
            x29_add(acc, ae, q, qe);
        }
    }
    x29_store(&sh[threadIdx.x], acc, ae);
    __syncthreads();
    x29_block_tree(sh);
    if (threadIdx.x == 0) {
        const X29 r = x29_load(&sh[0], ae);
        if (parts == 1)
            xyzz_store(slot_q(slots, g, W, nbits, bit), x29_to_xyzz(r, ae));
        else
            x29_store(&qpart[((size_t)g * (nbits + 1) + bit) * parts + part], r, ae);
    }
}

// the parts of each (item window, bit) sum: a workgroup takes 256 / parts
// sums, one thread per part, and adds them as a tree in LDS
__global__ __launch_bounds__(256) void k_msm_masked_merge(const X29S *__restrict__ qpart, uint32_t parts,
                                                          uint32_t nbits, uint32_t W, uint32_t gw,
                                                          uint8_t *const *__restrict__ slots) {
    __shared__ X29S sh[256];
    const uint32_t per = 256 / parts, z = threadIdx.x % parts;
    const uint64_t total = (uint64_t)gw * (nbits + 1);
    const uint64_t i = (uint64_t)blockIdx.x * per + threadIdx.x / parts;
    const bool live = i < total;
    bool ae = true, qe;
    X29 v;
    if (live) v = x29_load(&qpart[i * parts + z], ae);
    x29_store(&sh[threadIdx.x], v, ae);
    __syncthreads();
    for (uint32_t st = parts / 2; st; st >>= 1) {
        if (z < st) {
            X29 l = x29_load(&sh[threadIdx.x], ae);
            const X29 r = x29_load(&sh[threadIdx.x + st], qe);
            x29_add(l, ae, r, qe);
            x29_store(&sh[threadIdx.x], l, ae);
        }
        __syncthreads();
    }
    if (live && z == 0) {
        const X29 r = x29_load(&sh[threadIdx.x], ae);
        xyzz_store(slot_q(slots, (uint32_t)(i / (nbits + 1)), W, nbits, (uint32_t)(i % (nbits + 1))),
                   x29_to_xyzz(r, ae));
    }
}

struct SliceHeader {
    uint32_t magic, c, W, nbits, pad[4];
};

__host__ __device__ inline size_t slice_sums_off(uint32_t W, uint32_t nbits) {
    return kSliceQ + (size_t)W * (nbits + 1) * sizeof(G1Xyzz);
}
__host__ __device__ inline size_t slice_bytes_needed(uint32_t W, uint32_t nbits) {
    return slice_sums_off(W, nbits) + (size_t)W * sizeof(G1Jac);
}

__global__ void k_slice_header(uint8_t *const *slots, uint32_t count, uint32_t c, uint32_t W, uint32_t nbits) {
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= count) return;
    SliceHeader *h = reinterpret_cast<SliceHeader *>(slots[m] + 64);
    h->c = c;
    h->W = W;
    h->nbits = nbits;
    h->magic = kMagic;
}

__host__ __device__ __forceinline__ bool header_ok(const SliceHeader *h) {
    return h->magic == kMagic && h->W >= 1 && h->W <= 128 && h->c >= 2 && h->c <= kFixedMaxC && h->nbits <= 16;
}

// Device tail, step 1 (grid: W threads per slice, y = slice):
// S_w = A + L1 * sum_bit 2^bit Q_bit, written to the slice's S array.
__global__ __launch_bounds__(64) void k_msm_window_finalize(uint8_t *const *__restrict__ slices) {
    uint8_t *slice = slices[blockIdx.y];
    const SliceHeader *h = reinterpret_cast<const SliceHeader *>(slice + 64);
    if (!header_ok(h)) return;
    const uint32_t W = h->W, nbits = h->nbits;
    uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) return;
    const G1Xyzz *qw = reinterpret_cast<const G1Xyzz *>(slice + kSliceQ) + (size_t)w * (nbits + 1);
    G1Jac acc = xyzz_to_jac(nbits ? xyzz_load(&qw[nbits - 1]) : xyzz_id());
    for (int b = (int)nbits - 2; b >= 0; b--) {
        acc = jac_dbl(acc);
        acc = jac_add(acc, xyzz_to_jac(xyzz_load(&qw[b])));
    }
    for (uint32_t k = 1; k < kL1; k <<= 1) acc = jac_dbl(acc);
    acc = jac_add(acc, xyzz_to_jac(xyzz_load(&qw[nbits])));
    jac_store(reinterpret_cast<G1Jac *>(slice + slice_sums_off(W, nbits)) + w, acc);
}

// Device tail, step 2: Horner over windows; result (z = 1) at slice[0].
__device__ void collect_slice(uint8_t *slice) {
    const SliceHeader *h = reinterpret_cast<const SliceHeader *>(slice + 64);
    if (!header_ok(h)) {
        // no masked sums in this slice (msm() failed or was not run): write
        // the no-result marker (every byte 0xff: coordinates >= p, so no
        // point and not the identity) rather than reading garbage
        uint32_t *o = reinterpret_cast<uint32_t *>(slice);
        for (int i = 0; i < 24; i++) o[i] = 0xffffffffu;
        return;
    }
    const uint32_t c = h->c, W = h->W;
    const G1Jac *sums = reinterpret_cast<const G1Jac *>(slice + slice_sums_off(W, h->nbits));
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

// device-side slice header reset so a stale header is never trusted
__global__ void k_slice_clear(uint8_t *const *slots, uint32_t count) {
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < count) reinterpret_cast<SliceHeader *>(slots[m] + 64)->magic = 0;
}

// Fixed-base table T[w * n + i] = 2^(c w) P_i for w < Wd (affine, Montgomery,
// the identity kept as (0, 0)): one thread per point, c Jacobian doublings
// and one inversion per window.  A set-up pass (about 0.1 s at 2^22 points).
__global__ __launch_bounds__(256) void k_msm_table(const G1Aff *__restrict__ pts, G1Aff *__restrict__ table,
                                                   uint32_t n, uint32_t c, uint32_t Wd) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const G1Aff p = aff_load(&pts[i]);
    fe_store<FpParams>(&table[i].x, p.x);
    fe_store<FpParams>(&table[i].y, p.y);
    G1Jac j;
    j.x = p.x;
    j.y = p.y;
    j.z = aff_is_id(p) ? fe_zero<FpParams>() : fe_one<FpParams>();
    for (uint32_t w = 1; w < Wd; w++) {
        for (uint32_t k = 0; k < c; k++) j = jac_dbl(j);
        Fp x = fe_zero<FpParams>(), y = fe_zero<FpParams>();
        if (!fe_is_zero<FpParams>(j.z)) {
            const Fp zi = fe_inv<FpParams>(j.z);
            const Fp zi2 = fe_sqr<FpParams>(zi);
            x = j.x * zi2;
            y = j.y * zi2 * zi;
        }
        G1Aff *o = &table[(size_t)w * n + i];
        fe_store<FpParams>(&o->x, x);
        fe_store<FpParams>(&o->y, y);
    }
}

// Opt-in phase timing: when enabled (and the stream is not being captured), msm_run records a HIP event on the
// launch stream after each phase, waits for the last one and accumulates
// the per-phase durations (ms) and call counts.
enum MsmPhase { PH_COUNT = 0, PH_SCAN, PH_SCATTER, PH_ACCUMULATE, PH_FIXUP, PH_REDUCE, PH_HEADER, PH_NUM };
// Phase timing: each profiled call records its PH_NUM + 1 events into the
// next of kProfSets event sets and does not wait for them; the sets are read
// (synchronised and summed) by zk_msm_profile_read, or when a set is about to
// be reused.  So a profiled MSM is timed as an unprofiled one runs, with no
// host round trip added.
constexpr int kProfSets = 64;
struct PhaseProf {
    bool on = false;
    double ms[PH_NUM] = {0};
    uint64_t calls = 0;
    hipEvent_t ev[kProfSets][PH_NUM + 1] = {};
    int cur = 0;         // the set of the call in progress
    int pending = 0;     // recorded sets not yet summed: the last `pending` before cur
    bool init = false;
};
PhaseProf g_prof;

// sums set k into ms (waits for its last event)
hipError_t prof_collect(int k) {
    ZK_CHECK(hipEventSynchronize(g_prof.ev[k][PH_NUM]));
    for (int p = 0; p < PH_NUM; p++) {
        float ms = 0;
        ZK_CHECK(hipEventElapsedTime(&ms, g_prof.ev[k][p], g_prof.ev[k][p + 1]));
        g_prof.ms[p] += ms;
    }
    g_prof.calls++;
    return hipSuccess;
}
// every pending set, oldest first
hipError_t prof_drain() {
    while (g_prof.pending > 0) {
        const int k = ((g_prof.cur - g_prof.pending) % kProfSets + kProfSets) % kProfSets;
        g_prof.pending--;
        ZK_CHECK(prof_collect(k));
    }
    return hipSuccess;
}

hipError_t prof_mark(bool on, int k, hipStream_t stream) {
    if (!on) return hipSuccess;
    if (!g_prof.init) {
        for (auto &set : g_prof.ev)
            for (auto &e : set) ZK_CHECK(hipEventCreate(&e));
        g_prof.init = true;
    }
    if (k == 0 && g_prof.pending == kProfSets) {
        // the ring is full: the oldest set (the one about to be reused) first
        g_prof.pending--;
        ZK_CHECK(prof_collect(g_prof.cur));
    }
    return hipEventRecord(g_prof.ev[g_prof.cur][k], stream);
}
hipError_t prof_finish(bool on) {
    if (!on) return hipSuccess;
    g_prof.cur = (g_prof.cur + 1) % kProfSets;
    g_prof.pending++;
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

// One MSM of a batch on the host side: device point / scalar arrays and the
// curve slice that receives the masked sums and, after the tail, the result.
struct MsmJob {
    const G1Aff *pts;
    const Fr *scalars;
    uint32_t n;
    uint8_t *slice;
};

size_t msm_scratch_bytes(const uint32_t *ns, uint32_t count, uint32_t c, bool fixed = false) {
    return plan_total(make_plan(ns, count, c, fixed));
}
uint32_t msm_default_window(uint32_t n) { return default_window(n); }
// window for a batch: the default for its mean size
uint32_t batch_default_window(const uint32_t *ns, uint32_t count) {
    uint64_t t = 0;
    for (uint32_t m = 0; m < count; m++) t += ns[m];
    return default_window((uint32_t)(t / count ? t / count : 1));
}

// Picks the window and places the arrays: the requested window, or (c == 0)
// the default window first, then larger windows (fewer entries), then
// smaller ones, until the arrays fit the regions.  Fixed-base mode: only the
// table's window c.  Host-only.
bool choose_plan(const uint32_t *ns, uint32_t count, size_t slice_bytes, uint32_t c, uint8_t *const *regions,
                 const size_t *region_bytes, int nreg, MsmPlan &p, bool fixed = false) {
    const uint32_t c0 = c ? c : batch_default_window(ns, count);
    uint32_t order[32];
    int norder = 0;
    order[norder++] = c0;
    if (!c && !fixed) {
        for (uint32_t cc = c0 + 1; cc <= 17; cc++) order[norder++] = cc;
        for (uint32_t cc = c0 - 1; cc >= 2; cc--) order[norder++] = cc;
    }
    for (int k = 0; k < norder; k++) {
        // the small-pass chunk first, kChunk when its partials do not fit
        for (int small = 1; small >= 0; small--) {
            p = make_plan(ns, count, order[k], fixed, small != 0);
            // entry positions are u32 and the masked-sum grid's y is < 2^16
            if (p.max_entries < 0x80000000ull && (uint64_t)count * p.W < 65536 &&
                slice_bytes_needed(p.W, p.nbits) <= slice_bytes && place(p, regions, region_bytes, nreg))
                return true;
            if (p.chunk == kChunk) break;  // the same plan again
        }
    }
    return false;
}

// Runs a batch of MSMs up to their masked sums, written into each job's
// slice (header at slice + 64, Q at slice + 128; a slice needs
// slice_bytes_needed(W, nbits) <= slice_bytes).  All items go through every
// phase together: one launch per phase for the whole batch.  Scratch arrays
// are placed into the given regions; no allocation happens.  c == 0 picks the
// largest window <= the default that fits the regions.  *slots_dev receives
// the device array of slice pointers (inside the scratch) for a device tail.
// fixed: every job's points are a table built for window c (kFixedMinC <= c
// <= kFixedMaxC), see zk_msm_table_build.
hipError_t msm_run(const MsmJob *jobs, uint32_t count, size_t slice_bytes, uint32_t c, uint8_t *const *regions,
                   const size_t *region_bytes, int nreg, hipStream_t stream, uint8_t ***slots_dev,
                   bool fixed = false) {
    if (count == 0 || !jobs || nreg < 1 || nreg > 8) return hipErrorInvalidValue;
    if (fixed) {
        if (c < kFixedMinC || c > kFixedMaxC) return hipErrorInvalidValue;
    } else if (c != 0 && (c < 2 || c > 17)) {
        return hipErrorInvalidValue;  // sort kernels: <= 8 + 8 bucket bits
    }
    // item sizes: on the stack for small batches (no host allocation on the
    // path of a single MSM)
    uint32_t ns_small[64];
    std::vector<uint32_t> ns_big;
    uint32_t *ns = ns_small;
    if (count > 64) {
        ns_big.resize(count);
        ns = ns_big.data();
    }
    uint64_t ntot = 0;
    for (uint32_t m = 0; m < count; m++) {
        const MsmJob &j = jobs[m];
        if (j.n == 0 || j.n > 0x7fffffffu || !j.pts || !j.scalars || !j.slice) return hipErrorInvalidValue;
        ns[m] = j.n;
        ntot += j.n;
    }
    MsmPlan p;
    if (!choose_plan(ns, count, slice_bytes, c, regions, region_bytes, nreg, p, fixed)) return hipErrorOutOfMemory;
    c = p.c;
    // phase timing waits on events, which a stream under graph capture cannot
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    ZK_CHECK(hipStreamIsCapturing(stream, &cap));
    const bool prof = g_prof.on && cap == hipStreamCaptureStatusNone;
    MsmItemDev *items = (MsmItemDev *)p.ptr[A_ITEMS];
    uint8_t **slots = (uint8_t **)p.ptr[A_SLOTS];
    uint32_t *hist = (uint32_t *)p.ptr[A_HIST];
    uint32_t *offsets = (uint32_t *)p.ptr[A_OFFSETS];
    uint32_t *starts = (uint32_t *)p.ptr[A_STARTS];  // dead after k_msm_sort_group
    uint32_t *entries = (uint32_t *)p.ptr[A_ENTRIES];
    uint32_t *digits = entries;  // dead before the final entries are written
    void *tmp = p.ptr[A_TMP];  // dead after k_msm_sort_group
    X29S *buckets = (X29S *)p.ptr[A_BUCKETS];
    X29S *phead = (X29S *)p.ptr[A_PHEAD];
    X29S *ptail = (X29S *)p.ptr[A_PTAIL];
    uint32_t *hb = (uint32_t *)p.ptr[A_HB];
    uint32_t *tb = (uint32_t *)p.ptr[A_TB];
    X29S *ws1 = (X29S *)p.ptr[A_WS1];
    X29S *sum1 = (X29S *)p.ptr[A_SUM1];
    uint32_t *scan_sums = (uint32_t *)p.ptr[A_SCAN];
    uint32_t *longs = (uint32_t *)p.ptr[A_LONG];

    ZK_CHECK(prof_mark(prof, PH_COUNT, stream));
    // item table with prefixes, written from kernel arguments
    {
        uint32_t tile = 0;
        uint64_t dig = 0;
        for (uint32_t first = 0; first < count; first += kItemsPerLaunch) {
            ItemPack pk;
            memset(&pk, 0, sizeof(pk));
            pk.first = first;
            pk.k = count - first < kItemsPerLaunch ? count - first : kItemsPerLaunch;
            for (uint32_t j = 0; j < pk.k; j++) {
                const MsmJob &jb = jobs[first + j];
                MsmItemDev &it = pk.it[j];
                it.pts = jb.pts;
                it.scalars = jb.scalars;
                it.n = jb.n;
                it.tile0 = tile;
                it.ntiles = (jb.n + kSortTile - 1) / kSortTile;
                it.dig0 = dig;
                pk.slot[j] = jb.slice;
                tile += it.ntiles;
                dig += (uint64_t)jb.n * p.Wd;
            }
            k_msm_put_items<<<1, kItemsPerLaunch, 0, stream>>>(pk, items, slots, longs);
            ZK_CHECK(hipGetLastError());
        }
    }
    k_slice_clear<<<(count + 63) / 64, 64, 0, stream>>>(slots, count);
    ZK_CHECK(hipGetLastError());
    const uint32_t fx = p.fixed ? 1u : 0u;
    if (p.fixed) {
        if ((uint64_t)p.Wd * p.nbins > kHistMaxFixed || p.nbins > kFixedMaxBins || (1u << p.lo_bits) > kFixedMaxKeys)
            return hipErrorInvalidValue;  // not for kFixedMinC <= c <= kFixedMaxC
        k_msm_digits_hist<kHistMaxFixed><<<(uint32_t)p.ntiles, kDigitThreads, 0, stream>>>(
            items, count, c, p.Wd, p.lo_bits, p.nbins, fx, digits, hist);
    } else {
        if ((uint64_t)p.W * p.nbins > kHistMax || p.nbins > 256 || p.lo_bits > 8)
            return hipErrorInvalidValue;  // not for 2 <= c <= 17
        k_msm_digits_hist<kHistMax><<<(uint32_t)p.ntiles, kDigitThreads, 0, stream>>>(items, count, c, p.Wd, p.lo_bits,
                                                                                   p.nbins, fx, digits, hist);
    }
    ZK_CHECK(hipGetLastError());
    ZK_CHECK(prof_mark(prof, PH_SCAN, stream));
    ZK_CHECK(launch_scan(hist, (uint32_t)p.nhist, starts, nullptr, scan_sums, stream));
    ZK_CHECK(prof_mark(prof, PH_SCATTER, stream));
    const uint32_t ngroups = count * p.W * p.nbins;
    const dim3 sgrid((uint32_t)p.ntiles, p.Wd);
    if (p.fixed) {
        k_msm_scatter_hi<false, kFixedMaxBins><<<sgrid, 256, 0, stream>>>(items, count, p.Wd, digits, p.lo_bits,
                                                                         p.nbins, fx, hist, starts, (uint2 *)tmp);
        ZK_CHECK(hipGetLastError());
        k_msm_sort_group<false, kFixedMaxKeys><<<ngroups, kGroupThreads, 0, stream>>>(
            items, p.W, p.Wd, p.nbins, fx, (const uint2 *)tmp, starts, p.lo_bits, p.B, ngroups, entries, offsets);
    } else if (p.packed) {
        k_msm_scatter_hi<true, 256><<<sgrid, 256, 0, stream>>>(items, count, p.Wd, digits, p.lo_bits, p.nbins, fx,
                                                               hist, starts, (uint32_t *)tmp);
        ZK_CHECK(hipGetLastError());
        k_msm_sort_group<true, 256><<<ngroups, kGroupThreads, 0, stream>>>(
            items, p.W, p.Wd, p.nbins, fx, (const uint32_t *)tmp, starts, p.lo_bits, p.B, ngroups, entries, offsets);
    } else {
        k_msm_scatter_hi<false, 256><<<sgrid, 256, 0, stream>>>(items, count, p.Wd, digits, p.lo_bits, p.nbins, fx,
                                                                hist, starts, (uint2 *)tmp);
        ZK_CHECK(hipGetLastError());
        k_msm_sort_group<false, 256><<<ngroups, kGroupThreads, 0, stream>>>(
            items, p.W, p.Wd, p.nbins, fx, (const uint2 *)tmp, starts, p.lo_bits, p.B, ngroups, entries, offsets);
    }
    ZK_CHECK(hipGetLastError());
    ZK_CHECK(prof_mark(prof, PH_ACCUMULATE, stream));
    const uint32_t cblocks = (uint32_t)((p.chunks + 255) / 256);
    k_msm_accumulate<<<cblocks, 256, 0, stream>>>(entries, offsets, (uint32_t)p.NB, p.W * p.B, items, buckets,
                                                  phead, ptail, hb, tb, p.chunks, p.chunk);
    ZK_CHECK(hipGetLastError());
    ZK_CHECK(prof_mark(prof, PH_FIXUP, stream));
    k_msm_fixup<<<cblocks, 256, 0, stream>>>(buckets, phead, ptail, tb, offsets, longs, p.chunks, p.chunk);
    ZK_CHECK(hipGetLastError());
    {
        const uint64_t max_long = p.chunks / kWalk + 1;
        k_msm_fixup_long<<<(uint32_t)(max_long < 1024 ? max_long : 1024), 256, 0, stream>>>(buckets, phead, ptail,
                                                                                           longs);
        ZK_CHECK(hipGetLastError());
    }
    ZK_CHECK(prof_mark(prof, PH_REDUCE, stream));
    const uint32_t gw = count * p.W;
    const uint32_t nseg = gw * p.T1;
    if (p.T1 >= 64) {
        // one wave per 64 segments; its 8 partials go to the ws1 array (gw *
        // T1 / 8 entries, within its gw * T1)
        k_msm_reduce_fold<<<nseg / 64, 64, 0, stream>>>(buckets, offsets, p.B, p.T1, ws1);
        ZK_CHECK(hipGetLastError());
        k_msm_fold_sums<<<dim3(p.nbits + 1, gw), 64, 0, stream>>>(ws1, p.T1, p.nbits, p.W, slots);
        ZK_CHECK(hipGetLastError());
    } else {
        k_msm_reduce_l1<<<(nseg + 63) / 64, 64, 0, stream>>>(buckets, offsets, gw, p.B, p.T1, ws1, sum1);
        ZK_CHECK(hipGetLastError());
        // split each masked sum over `parts` workgroups when there are too
        // few sums to give every SIMD a wave (the sums are latency-bound
        // chains of point additions; more parts than that only adds tree
        // levels); the parts go to the chunk partials' array (dead after the
        // fixup)
        const uint64_t sums = (uint64_t)gw * (p.nbits + 1);
        uint32_t parts = 1;
        while (sums * parts * 2 <= 1024 && parts * 2 * 256 <= p.T1 && parts < 256 && sums * parts * 2 <= p.chunks)
            parts *= 2;
        k_msm_masked_sums<<<dim3(p.nbits + 1, gw, parts), 256, 0, stream>>>(ws1, sum1, p.T1, p.nbits, p.W, slots,
                                                                          phead);
        ZK_CHECK(hipGetLastError());
        if (parts > 1) {
            const uint32_t per = 256 / parts;
            k_msm_masked_merge<<<(uint32_t)((sums + per - 1) / per), 256, 0, stream>>>(phead, parts, p.nbits, p.W,
                                                                                      gw, slots);
            ZK_CHECK(hipGetLastError());
        }
    }
    ZK_CHECK(prof_mark(prof, PH_HEADER, stream));
    k_slice_header<<<(count + 63) / 64, 64, 0, stream>>>(slots, count, c, p.W, p.nbits);
    ZK_CHECK(hipGetLastError());
    ZK_CHECK(prof_mark(prof, PH_NUM, stream));
    if (slots_dev) *slots_dev = slots;
    return prof_finish(prof);
}

// Tail on the host: one Horner pass over the exponents of the masked sums
// (msm_tail in host_ec.h), a serial chain of ~250 XYZZ doublings and ~210
// additions that a single GPU wave would run in milliseconds; a host core
// runs it in ~0.25 ms.  The Q points are copied down and the normalised
// result is written back into each slice.  split: the exponents below ~0.55
// of the top go to a second thread (their sum is added after the main chain
// has doubled its part down to them), for a lone MSM with idle cores.
// The slice's header has been checked by the caller (header_ok).
void host_tail_one(const uint8_t *blob, uint8_t *out96, bool split = false) {
    const SliceHeader *h = reinterpret_cast<const SliceHeader *>(blob + 64);
    static_assert(sizeof(zkh::Xyzz) == sizeof(G1Xyzz), "host and device XYZZ layouts differ");
    static_assert((kL1 & (kL1 - 1)) == 0, "kL1 is a power of two");
    const zkh::Xyzz *q = reinterpret_cast<const zkh::Xyzz *>(blob + kSliceQ);
    const uint32_t c = h->c, W = h->W, nb = h->nbits, l1 = (uint32_t)__builtin_ctz(kL1);
    const int top = zkh::msm_tail_top(c, W, nb, l1), lo = top * 11 / 20;
    zkh::Jac res;
    if (!split || lo < 1) {
        res = zkh::msm_tail(q, c, W, nb, l1);
    } else {
        zkh::Xyzz low;
        auto low_part = [&]() { low = zkh::msm_tail_range(q, c, W, nb, l1, 0, lo - 1); };
        std::thread th;
        try {
            th = std::thread(low_part);
        } catch (...) {
        }
        zkh::Xyzz a = zkh::msm_tail_range(q, c, W, nb, l1, lo, top);
        for (int k = 0; k < lo; k++) a = zkh::xyzz_dbl(a);
        if (th.joinable())
            th.join();
        else
            low_part();  // no thread could be started
        res = zkh::xyzz_normalize(zkh::xyzz_add(a, low));
    }
    memcpy(out96, &res, 96);
}

// Pinned host staging for the collects' copies (device-to-host of the masked
// sums, host-to-device of the results): buffers from hipHostMalloc, taken
// from a free list per call and returned to it, kept for the process's life,
// so concurrent collects each stage in their own and small copies go by
// direct DMA instead of the runtime's pageable staging.  A null buffer (no
// pinned memory to be had) makes the caller stage in pageable memory.
struct Stage {
    uint8_t *p = nullptr;
    size_t cap = 0;
};
std::mutex g_stage_mu;
std::vector<Stage> g_stage_free;

Stage stage_take(size_t need) {
    {
        std::lock_guard<std::mutex> g(g_stage_mu);
        for (size_t i = 0; i < g_stage_free.size(); i++)
            if (g_stage_free[i].cap >= need) {
                const Stage st = g_stage_free[i];
                g_stage_free.erase(g_stage_free.begin() + (long)i);
                return st;
            }
    }
    size_t cap = (size_t)1 << 16;
    while (cap < need) cap <<= 1;
    void *p = nullptr;
    if (hipHostMalloc(&p, cap, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();  // not left for the next call to report
        return Stage{};
    }
    return Stage{static_cast<uint8_t *>(p), cap};
}
// returns the buffer to the free list (freed if the list cannot grow)
struct StageHold {
    Stage st;
    std::vector<uint8_t> fallback;
    // the staging area: pinned when available, else pageable
    uint8_t *take(size_t need) {
        st = stage_take(need);
        if (st.p) return st.p;
        fallback.assign(need, 0);
        return fallback.data();
    }
    ~StageHold() {
        if (!st.p) return;
        try {
            std::lock_guard<std::mutex> g(g_stage_mu);
            g_stage_free.push_back(st);
        } catch (...) {
            (void)hipHostFree(st.p);
        }
    }
};

// the host tails of `batch` slice images at img + m * stride into out + m * 96,
// over up to 16 threads (a lone MSM splits its own tail in two)
void host_tails(const uint8_t *img, size_t stride, uint32_t batch, uint8_t *out) {
    unsigned hw = std::thread::hardware_concurrency();
    unsigned nt = batch < 2 ? 1 : (hw ? (hw < batch ? hw : batch) : 1);
    if (nt > 16) nt = 16;
    auto run = [&](unsigned t) {
        for (uint32_t m = t; m < batch; m += nt) host_tail_one(img + (size_t)m * stride, out + (size_t)m * 96, batch == 1);
    };
    if (nt <= 1) {
        run(0);
        return;
    }
    std::vector<std::thread> th;
    unsigned started = 0;
    try {
        for (unsigned t = 0; t < nt; t++, started++) th.emplace_back(run, t);
    } catch (...) {
    }
    for (auto &x : th) x.join();
    for (unsigned t = started; t < nt; t++) run(t);  // shares whose thread could not be started
}

// slice_bytes (optional): every slice holds at least this many bytes, so
// one copy of min(slice_bytes, kSliceReserve) per slice brings its header and
// masked sums together (a valid header's sums always fit: msm() placed them
// within that bound), saving the header-first round trip.
hipError_t host_collect(uint8_t *const *slices, uint32_t batch, hipStream_t stream, size_t slice_bytes = 0) {
    if (slice_bytes >= kSliceQ) {
        const size_t bytes = slice_bytes < kSliceReserve ? slice_bytes : kSliceReserve;
        StageHold hold;
        uint8_t *img = hold.take((size_t)batch * (bytes + 96));  // slice images, then the results
        uint8_t *out = img + (size_t)batch * bytes;
        for (uint32_t m = 0; m < batch; m++)
            ZK_CHECK(hipMemcpyAsync(img + (size_t)m * bytes, slices[m], bytes, hipMemcpyDeviceToHost, stream));
        ZK_CHECK(hipStreamSynchronize(stream));
        for (uint32_t m = 0; m < batch; m++) {
            const SliceHeader *h = reinterpret_cast<const SliceHeader *>(img + (size_t)m * bytes + 64);
            // no result (cleared or stale header), or sums beyond what was read
            if (!header_ok(h) || slice_sums_off(h->W, h->nbits) > bytes) return hipErrorInvalidValue;
        }
        host_tails(img, bytes, batch, out);
        for (uint32_t m = 0; m < batch; m++)
            ZK_CHECK(hipMemcpyAsync(slices[m], out + (size_t)m * 96, 96, hipMemcpyHostToDevice, stream));
        return hipStreamSynchronize(stream);  // the staging must outlive the copies
    }
    // the largest slice blob: header + Q for W <= 128, nbits <= 16 is bounded
    // by slice_sums_off(W, nbits); read the header first to size each copy
    std::vector<SliceHeader> hdr(batch);
    for (uint32_t m = 0; m < batch; m++)
        ZK_CHECK(hipMemcpyAsync(&hdr[m], slices[m] + 64, sizeof(SliceHeader), hipMemcpyDeviceToHost, stream));
    ZK_CHECK(hipStreamSynchronize(stream));
    // a slice msm() never finished (cleared or stale header) has no result:
    // an error, not a commitment
    for (uint32_t m = 0; m < batch; m++)
        if (!header_ok(&hdr[m])) return hipErrorInvalidValue;
    size_t stride = 0;  // the largest slice image: each is staged at m * stride
    for (uint32_t m = 0; m < batch; m++) {
        const size_t bytes = slice_sums_off(hdr[m].W, hdr[m].nbits);
        if (bytes > stride) stride = bytes;
    }
    StageHold hold;
    uint8_t *img = hold.take((size_t)batch * (stride + 96));  // slice images, then the results
    uint8_t *out = img + (size_t)batch * stride;
    for (uint32_t m = 0; m < batch; m++)
        ZK_CHECK(hipMemcpyAsync(img + (size_t)m * stride, slices[m], slice_sums_off(hdr[m].W, hdr[m].nbits),
                                hipMemcpyDeviceToHost, stream));
    ZK_CHECK(hipStreamSynchronize(stream));
    host_tails(img, stride, batch, out);
    for (uint32_t m = 0; m < batch; m++)
        ZK_CHECK(hipMemcpyAsync(slices[m], out + (size_t)m * 96, 96, hipMemcpyHostToDevice, stream));
    return hipStreamSynchronize(stream);  // the staging must outlive the copies
}

// host_collect for slices at base + m * stride (stride >= kSliceReserve, the
// layout zk_msm_batch uses): one device-to-host copy of the whole region, the
// tails on the host, one host-to-device copy back, instead of three small
// copies per slice.  jac_out (optional, host): the count results, 96 B each.
hipError_t host_collect_strided(uint8_t *base, size_t stride, uint32_t batch, hipStream_t stream,
                                uint8_t *jac_out) {
    if (stride < kSliceReserve) return hipErrorInvalidValue;
    const size_t bytes = (size_t)(batch - 1) * stride + kSliceReserve;
    StageHold hold;
    uint8_t *img = hold.take(bytes + (size_t)batch * 96);  // the region's image, then the results
    uint8_t *out = img + bytes;
    ZK_CHECK(hipMemcpyAsync(img, base, bytes, hipMemcpyDeviceToHost, stream));
    ZK_CHECK(hipStreamSynchronize(stream));
    // a slice msm() never finished (cleared or stale header) has no result
    for (uint32_t m = 0; m < batch; m++)
        if (!header_ok(reinterpret_cast<const SliceHeader *>(img + (size_t)m * stride + 64))) return hipErrorInvalidValue;
    host_tails(img, stride, batch, out);
    for (uint32_t m = 0; m < batch; m++) memcpy(img + (size_t)m * stride, out + (size_t)m * 96, 96);
    ZK_CHECK(hipMemcpyAsync(base, img, bytes, hipMemcpyHostToDevice, stream));
    if (jac_out) memcpy(jac_out, out, (size_t)batch * 96);
    return hipStreamSynchronize(stream);  // the staging must outlive the copy
}

bool device_tail_requested() {
    const char *e = getenv("ZK_MSM_TAIL");
    return e && strcmp(e, "device") == 0;
}

hipError_t device_collect(uint8_t *const *slices_dev, uint32_t batch, hipStream_t stream) {
    // W <= 128: two blocks of 64 threads per slice cover every window
    k_msm_window_finalize<<<dim3(2, batch), 64, 0, stream>>>(slices_dev);
    ZK_CHECK(hipGetLastError());
    k_msm_collect<<<(batch + 63) / 64, 64, 0, stream>>>(slices_dev, batch);
    return hipGetLastError();
}

// One MSM enqueued on stream with the device tail: no host wait and no host
// allocation; the result lands in out[0..96) once the stream reaches it
// (96 bytes of 0xff if the MSM did not finish).  out holds ZK_MSM_OUT_BYTES.
hipError_t msm_one_async(void *out, const void *points, const void *scalars, uint32_t n, uint32_t c, void *scratch,
                         size_t scratch_bytes, hipStream_t stream) {
    uint8_t *regions[1] = {(uint8_t *)scratch};
    size_t bytes[1] = {scratch_bytes};
    MsmJob job{(const G1Aff *)points, (const Fr *)scalars, n, (uint8_t *)out};
    uint8_t **slots = nullptr;
    ZK_CHECK(msm_run(&job, 1, ZK_MSM_OUT_BYTES, c, regions, bytes, 1, stream, &slots));
    return device_collect(slots, 1, stream);
}

// Items of a batch by kind: plain points (table_c = 0) and, per table
// window, fixed-base tables; each kind goes through the pipeline as one pass.
// Plain first, then tables by window.  False if an item is invalid.
struct ItemGroup {
    uint32_t table_c;
    std::vector<uint32_t> idx;
};
bool group_items(const zk_msm_item *items, int count, std::vector<ItemGroup> &groups) {
    groups.clear();
    for (int m = 0; m < count; m++) {
        const int tc = items[m].table_c;
        if (items[m].n <= 0) return false;
        if (tc != 0 && (tc < (int)kFixedMinC || tc > (int)kFixedMaxC)) return false;
        if (tc != 0 && (uint64_t)fixed_windows((uint32_t)tc) * (uint64_t)items[m].n >= 0x80000000ull) return false;
        size_t g = 0;
        while (g < groups.size() && groups[g].table_c != (uint32_t)tc) g++;
        if (g == groups.size()) groups.push_back(ItemGroup{(uint32_t)tc, {}});
        groups[g].idx.push_back((uint32_t)m);
    }
    std::sort(groups.begin(), groups.end(),
              [](const ItemGroup &x, const ItemGroup &y) { return x.table_c < y.table_c; });
    return true;
}

// scratch bytes for the batch: the largest pass (the passes run one after
// another in the same scratch); 0 if an item is invalid
size_t items_scratch_bytes(const zk_msm_item *items, int count, uint32_t c) {
    std::vector<ItemGroup> groups;
    if (count <= 0 || !items || !group_items(items, count, groups)) return 0;
    size_t best = 0;
    for (const ItemGroup &g : groups) {
        std::vector<uint32_t> ns(g.idx.size());
        for (size_t k = 0; k < g.idx.size(); k++) ns[k] = (uint32_t)items[g.idx[k]].n;
        const uint32_t cnt = (uint32_t)ns.size();
        const size_t b = g.table_c ? msm_scratch_bytes(ns.data(), cnt, g.table_c, true)
                                   : msm_scratch_bytes(ns.data(), cnt, c ? c : batch_default_window(ns.data(), cnt));
        best = b > best ? b : best;
    }
    return best + 256;
}

// Runs every pass of the batch up to the masked sums in item m's slice at out
// + m * ZK_MSM_OUT_BYTES; device_tail: each pass's results are combined on the
// device right after it (the pass's slice pointers live in the scratch, which
// the next pass reuses).  c: the window of the plain items (0 = default).
hipError_t run_items(void *out, const zk_msm_item *items, int count, uint32_t c, void *scratch, size_t scratch_bytes,
                     hipStream_t stream, bool device_tail) {
    std::vector<ItemGroup> groups;
    if (count <= 0 || !items || !out || !scratch || !group_items(items, count, groups)) return hipErrorInvalidValue;
    uint8_t *regions[1] = {(uint8_t *)scratch};
    size_t bytes[1] = {scratch_bytes};
    for (const ItemGroup &g : groups) {
        std::vector<MsmJob> jobs(g.idx.size());
        for (size_t k = 0; k < g.idx.size(); k++) {
            const zk_msm_item &it = items[g.idx[k]];
            jobs[k] = MsmJob{(const G1Aff *)it.points, (const Fr *)it.scalars, (uint32_t)it.n,
                             (uint8_t *)out + (size_t)g.idx[k] * ZK_MSM_OUT_BYTES};
        }
        uint8_t **slots = nullptr;
        ZK_CHECK(msm_run(jobs.data(), (uint32_t)jobs.size(), ZK_MSM_OUT_BYTES, g.table_c ? g.table_c : c, regions,
                         bytes, 1, stream, &slots, g.table_c != 0));
        if (device_tail) ZK_CHECK(device_collect(slots, (uint32_t)jobs.size(), stream));
    }
    return hipSuccess;
}

// zk_msm_batch whose results also land in host memory (jac_host: count x 96
// B) without a copy per item (the prover core's commitments)
hipError_t msm_batch_host(void *out, const zk_msm_item *items, int count, void *scratch, size_t scratch_bytes,
                          hipStream_t stream, uint8_t *jac_host) {
    if (count <= 0 || !items || !out || !scratch || !jac_host) return hipErrorInvalidValue;
    const bool dev = device_tail_requested();
    ZK_CHECK(run_items(out, items, count, 0, scratch, scratch_bytes, stream, dev));
    if (!dev) return host_collect_strided((uint8_t *)out, ZK_MSM_OUT_BYTES, (uint32_t)count, stream, jac_host);
    for (int m = 0; m < count; m++)
        ZK_CHECK(hipMemcpyAsync(jac_host + (size_t)m * 96, (uint8_t *)out + (size_t)m * ZK_MSM_OUT_BYTES, 96,
                                hipMemcpyDeviceToHost, stream));
    ZK_CHECK(hipStreamSynchronize(stream));
    for (int m = 0; m < count; m++) {  // the device tail's no-result marker: 96 bytes of 0xff
        const uint8_t *q = jac_host + (size_t)m * 96;
        bool all = true;
        for (int b = 0; b < 96; b++) all &= q[b] == 0xff;
        if (all) return hipErrorInvalidValue;
    }
    return hipSuccess;
}

// slice_bytes: as host_collect (0 = unknown)
hipError_t msm_collect(uint8_t *const *slices_dev, uint32_t batch, hipStream_t stream, size_t slice_bytes = 0) {
    if (!batch) return hipSuccess;
    if (device_tail_requested()) return device_collect(slices_dev, batch, stream);
    std::vector<uint8_t *> slices(batch);
    ZK_CHECK(hipMemcpyAsync(slices.data(), slices_dev, (size_t)batch * sizeof(uint8_t *), hipMemcpyDeviceToHost,
                            stream));
    ZK_CHECK(hipStreamSynchronize(stream));
    return host_collect(slices.data(), batch, stream, slice_bytes);
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
    return zk::guard([&]() -> hipError_t {
        (void)acc_indices_buf;
        (void)threads;
        (void)prepared_sort_indices_temp_storage_bytes;
        (void)skip_zero;  // zero digits never reach a bucket here; the result is the same
        if (array_len <= 0 || windows <= 0 || window_bits <= 0 || window_bits > 30 || workers <= 0)
            return hipErrorInvalidValue;
        size_t idx_bytes = (size_t)array_len * (size_t)windows * 4;
        size_t slice_bytes = (((size_t)windows << window_bits) + 2 * (size_t)workers) * sizeof(G1Jac);
        // the curve slice beyond the first kSliceReserve bytes (header, masked and
        // window sums) is free scratch as well: at least 2^10 * 26 Jacobian points
        uint8_t *slice_free = slice_bytes > kSliceReserve ? (uint8_t *)res + kSliceReserve : nullptr;
        uint8_t *regions[6] = {(uint8_t *)sort_buf, (uint8_t *)indices_buf0, (uint8_t *)indices_buf1,
                               (uint8_t *)indices_buf2, (uint8_t *)indices_buf3, slice_free};
        size_t bytes[6] = {3 * idx_bytes, idx_bytes, idx_bytes, idx_bytes, idx_bytes,
                           slice_free ? slice_bytes - kSliceReserve : 0};
        MsmJob job{(const G1Aff *)points, (const Fr *)scalars, (uint32_t)array_len, (uint8_t *)res};
        return msm_run(&job, 1, slice_bytes < kSliceReserve ? slice_bytes : kSliceReserve, 0, regions, bytes, 6,
                       stream, nullptr);
    });
}

// cuda/msm.cu:434-467 (caller src/cuda/msm.rs:369-455).  buckets is a device
// array of batch_size device pointers to the curve slices given to msm();
// afterwards each slice starts with that MSM's result (Jacobian, z = 1, or
// the identity (0, 1, 0)).  The remain/next_remain arrays are unused.
hipError_t batch_msm_collect(void *remain_indices, void *remain_acc, void *next_remain_indices,
                             void *next_remain_acc, void *buckets, unsigned workers, unsigned windows,
                             unsigned window_bits, unsigned batch_size, hipStream_t stream) {
    return zk::guard([&]() -> hipError_t {
        (void)remain_indices;
        (void)remain_acc;
        (void)next_remain_indices;
        (void)next_remain_acc;
        if (!buckets && batch_size) return hipErrorInvalidValue;
        // each slice is ((windows << window_bits) + 2 workers) Jacobian points
        // (src/cuda/msm.rs:297), as the caller sized it for msm()
        const uint64_t pts = window_bits < 40 ? ((uint64_t)windows << window_bits) + 2ull * workers : 0;
        const size_t slice_bytes = pts && pts < (1ull << 40) ? (size_t)pts * sizeof(G1Jac) : 0;
        return msm_collect((uint8_t *const *)buckets, batch_size, stream, slice_bytes);
    });
}

// Phase timing (bench instrumentation): enable / read / reset.  times[k]
// receives the accumulated ms of phase k (digits+histogram, scan, scatter,
// accumulate, fixup, reduce, header); returns the number of profiled calls.
void zk_msm_profile_enable(int on) { g_prof.on = on != 0; }
int zk_msm_profile_read(double *times, int ntimes, int reset) {
    if (prof_drain() != hipSuccess) return -1;
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
    uint32_t un = (uint32_t)n;
    uint32_t cc = c > 0 ? (uint32_t)c : default_window(un);
    return msm_scratch_bytes(&un, 1, cc) + 256;
}

// after a device tail: a slice without masked sums carries the no-result
// marker (every byte 0xff), which makes the call fail
static hipError_t check_markers(uint8_t *const *slices, uint32_t count, hipStream_t stream) {
    std::vector<uint32_t> z((size_t)count * 8);
    for (uint32_t m = 0; m < count; m++)
        ZK_CHECK(hipMemcpyAsync(&z[(size_t)m * 8], slices[m] + 64, 32, hipMemcpyDeviceToHost, stream));
    ZK_CHECK(hipStreamSynchronize(stream));
    for (uint32_t m = 0; m < count; m++)
        if (z[(size_t)m * 8] == 0xffffffffu && z[(size_t)m * 8 + 7] == 0xffffffffu) return hipErrorInvalidValue;
    return hipSuccess;
}

// Combines the masked sums of the given slices into their results (host or,
// with ZK_MSM_TAIL=device, device tail) and waits for them.
static hipError_t finish_batch(uint8_t *const *slices, uint8_t **slots_dev, uint32_t count, hipStream_t stream) {
    if (device_tail_requested()) {
        ZK_CHECK(device_collect(slots_dev, count, stream));
        return check_markers(slices, count, stream);
    }
    return host_collect(slices, count, stream);
}

// One MSM with caller-owned scratch (no allocation): the result (Jacobian,
// z = 1, or the identity (0, 1, 0)) lands in out[0..96); out must hold
// ZK_MSM_OUT_BYTES (result, header, masked and window sums).  c = 0 picks
// the default window for n.  Returns after the result is written.
hipError_t zk_msm(void *out, const void *points, const void *scalars, int n, int c, void *scratch,
                  size_t scratch_bytes, hipStream_t stream) {
    return zk::guard([&]() -> hipError_t {
        if (n <= 0 || !out || !scratch) return hipErrorInvalidValue;
        uint8_t *regions[1] = {(uint8_t *)scratch};
        size_t bytes[1] = {scratch_bytes};
        uint8_t *slice = (uint8_t *)out;
        MsmJob job{(const G1Aff *)points, (const Fr *)scalars, (uint32_t)n, slice};
        uint8_t **slots = nullptr;
        ZK_CHECK(msm_run(&job, 1, ZK_MSM_OUT_BYTES, c > 0 ? (uint32_t)c : 0, regions, bytes, 1, stream, &slots));
        return finish_batch(&slice, slots, 1, stream);
    });
}

// Scratch bytes zk_msm_batch needs for these items at window c (0: default;
// table items use their table's window).
size_t zk_msm_batch_scratch_bytes(const zk_msm_item *items, int count, int c) {
    return items_scratch_bytes(items, count, c > 0 ? (uint32_t)c : 0);
}

// count MSMs in one pass of the pipeline per kind of item (plain points, and
// tables of each window; one launch per phase for all items of a kind).
// items is a host array; item m's result lands in out + m * ZK_MSM_OUT_BYTES
// (Jacobian, z = 1).  Returns after the results are written.
hipError_t zk_msm_batch(void *out, const zk_msm_item *items, int count, int c, void *scratch,
                        size_t scratch_bytes, hipStream_t stream) {
    return zk::guard([&]() -> hipError_t {
        if (count <= 0 || !items || !out || !scratch) return hipErrorInvalidValue;
        const bool dev = device_tail_requested();
        ZK_CHECK(run_items(out, items, count, c > 0 ? (uint32_t)c : 0, scratch, scratch_bytes, stream, dev));
        if (!dev) return host_collect_strided((uint8_t *)out, ZK_MSM_OUT_BYTES, (uint32_t)count, stream, nullptr);
        std::vector<uint8_t *> slices(count);
        for (int m = 0; m < count; m++) slices[m] = (uint8_t *)out + (size_t)m * ZK_MSM_OUT_BYTES;
        return check_markers(slices.data(), (uint32_t)count, stream);
    });
}

// The window msm() would use with the reference driver's buffers for
// array_len points (src/cuda/msm.rs:284-311 sizes), or 0 if its scratch
// arrays would not fit.  Host-only: no device memory is touched.
int zk_msm_plan_window(int array_len, int windows, int window_bits, int workers) {
    if (array_len <= 0 || windows <= 0 || window_bits <= 0 || window_bits > 30 || workers <= 0) return 0;
    size_t idx_bytes = (size_t)array_len * (size_t)windows * 4;
    size_t slice_bytes = (((size_t)windows << window_bits) + 2 * (size_t)workers) * sizeof(G1Jac);
    // stand-in bases 2^40 apart: placement only does address arithmetic
    uint8_t *regions[6];
    for (int r = 0; r < 6; r++) regions[r] = (uint8_t *)((uintptr_t)(r + 1) << 40);
    size_t bytes[6] = {3 * idx_bytes, idx_bytes, idx_bytes, idx_bytes, idx_bytes,
                       slice_bytes > kSliceReserve ? slice_bytes - kSliceReserve : 0};
    uint32_t n = (uint32_t)array_len;
    MsmPlan p;
    if (!choose_plan(&n, 1, slice_bytes < kSliceReserve ? slice_bytes : kSliceReserve, 0, regions, bytes, 6, p))
        return 0;
    return (int)p.c;
}

// Window table access (a tuned table saved by the caller can be restored).
hipError_t zk_msm_set_window(int log_n, int c) {
    if (log_n < 0 || log_n > 31 || c < 0 || c == 1 || c > 17) return hipErrorInvalidValue;
    g_window[log_n].store((uint8_t)c, std::memory_order_relaxed);
    return hipSuccess;
}
int zk_msm_get_window(int log_n) {
    if (log_n < 0 || log_n > 31) return 0;
    return (int)default_window(log_n >= 31 ? 0x80000000u : 1u << log_n);
}

// Autotuner: times zk_msm on the caller's points / scalars (n of each,
// device) for windows around the model's choice and records the fastest for
// the size class ceil(log2 n).  ms (optional, 18 floats) receives the time
// per window (0 for windows not tried).  Allocates its own scratch.
hipError_t zk_msm_autotune(const void *points, const void *scalars, int n, int reps, hipStream_t stream, int *best,
                           float *ms) {
    return zk::guard([&]() -> hipError_t {
        if (n <= 0 || !points || !scalars) return hipErrorInvalidValue;
        if (reps < 1) reps = 1;
        const uint32_t lg = log2_exact((uint32_t)n);
        const uint8_t saved = lg < 32 ? g_window[lg].load(std::memory_order_relaxed) : 0;
        if (lg < 32) g_window[lg].store(0, std::memory_order_relaxed);
        const int c0 = (int)default_window((uint32_t)n);
        int lo = c0 - 2 < 4 ? 4 : c0 - 2, hi = c0 + 2 > 17 ? 17 : c0 + 2;
        size_t need = 0;
        for (int c = lo; c <= hi; c++) {
            size_t b = zk_msm_scratch_bytes(n, c);
            if (b > need) need = b;
        }
        void *scratch = nullptr, *out = nullptr;
        hipError_t err = hipMalloc(&scratch, need);
        if (err == hipSuccess) err = hipMalloc(&out, ZK_MSM_OUT_BYTES);
        float best_ms = 0;
        int best_c = c0;
        if (ms)
            for (int c = 0; c < 18; c++) ms[c] = 0;
        for (int c = lo; c <= hi && err == hipSuccess; c++) {
            err = zk_msm(out, points, scalars, n, c, scratch, need, stream);  // warm-up
            double t = 1e30;
            for (int r = 0; r < reps && err == hipSuccess; r++) {
                auto t0 = std::chrono::steady_clock::now();
                err = zk_msm(out, points, scalars, n, c, scratch, need, stream);
                double dt = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                if (dt < t) t = dt;
            }
            if (err != hipSuccess) break;
            if (ms) ms[c] = (float)t;
            if (best_ms == 0 || t < best_ms) {
                best_ms = (float)t;
                best_c = c;
            }
        }
        if (scratch) (void)hipFree(scratch);
        if (out) (void)hipFree(out);
        if (err != hipSuccess) {
            if (lg < 32) g_window[lg].store(saved, std::memory_order_relaxed);
            return err;
        }
        if (lg < 32) g_window[lg].store((uint8_t)best_c, std::memory_order_relaxed);
        if (best) *best = best_c;
        return hipSuccess;
    });
}

// zk_msm_batch without any host wait: the window combination runs on the
// device (one thread per item), so the call only enqueues kernels on stream
// and can be captured into a HIP graph.  Results land as in zk_msm_batch once
// the stream reaches them.  items is read during the call only.
hipError_t zk_msm_batch_async(void *out, const zk_msm_item *items, int count, int c, void *scratch,
                              size_t scratch_bytes, hipStream_t stream) {
    return zk::guard([&]() -> hipError_t {
        if (count <= 0 || !items || !out || !scratch) return hipErrorInvalidValue;
        return run_items(out, items, count, c > 0 ? (uint32_t)c : 0, scratch, scratch_bytes, stream, true);
    });
}

// Rows of a fixed-base table: ceil(255 / c) per point.
size_t zk_msm_table_bytes(int n, int c) {
    if (n <= 0 || c < (int)kFixedMinC || c > (int)kFixedMaxC) return 0;
    const uint64_t rows = (uint64_t)fixed_windows((uint32_t)c) * (uint64_t)n;
    return rows < 0x80000000ull ? (size_t)rows * sizeof(G1Aff) : 0;
}

// The table window for n points: fewest additions, ceil(255 / c) n to fill
// the buckets plus about 2^(c+1) to reduce them (the larger c on a tie), up
// to c = 17.  Wider windows need more than 8 + 8 sorted bucket bits, whose
// two sort levels then write shorter runs over fewer groups; measured on
// MI355X that costs more than the additions saved (2^20 points, batches of
// 8: 1.72 ms per MSM at c = 17, 1.84 ms at c = 20).
int zk_msm_table_window(int n) {
    if (n <= 0) return 0;
    int best = 0;
    double cost = 0;
    for (uint32_t c = kFixedMinC; c <= kFixedModelMaxC; c++) {
        if (!zk_msm_table_bytes(n, (int)c)) continue;
        const double t = (double)fixed_windows(c) * n + (double)(2u << c);
        if (!best || t <= cost) {
            best = (int)c;
            cost = t;
        }
    }
    return best;
}

hipError_t zk_msm_table_build(void *table, const void *points, int n, int c, hipStream_t stream) {
    return zk::guard([&]() -> hipError_t {
        if (!table || !points || !zk_msm_table_bytes(n, c)) return hipErrorInvalidValue;
        k_msm_table<<<((uint32_t)n + 255) / 256, 256, 0, stream>>>((const G1Aff *)points, (G1Aff *)table,
                                                                     (uint32_t)n, (uint32_t)c,
                                                                     fixed_windows((uint32_t)c));
        return hipGetLastError();
    });
}

}  // extern "C"
