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
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <vector>

#include "ff29.h"
#include "guard.h"
#include "zkcore.h"
#include "zkinternal.h"

using namespace zk;

namespace {

constexpr uint32_t kTile = 1024;
constexpr uint32_t kThreads = 256;
constexpr uint32_t kMaxDeg = 8;

// One pass of one NTT as a launch sees it.  A launch carries up to
// kJobsPerLaunch jobs (one per NTT of a batch) in its kernel arguments; job j
// owns the blocks [block0_j, block0_{j+1}).
struct NttJob {
    const Fr *x;
    Fr *y;
    const Fr *omegas;
    const Fr *n_inv;  // null except on an iNTT's last pass
    // first pass of a coset extension: x holds nz coefficients (the rest of
    // the 2^log_n input is zero and never read), element i scaled by
    // coset[(i % cn) - 1] when i % cn != 0; nz = 0: x is the whole input
    const Fr *coset;
    uint32_t nz, cn;
    uint32_t block0;
    uint8_t log_n, log_p, deg, log_g;
};
constexpr uint32_t kJobsPerLaunch = 48;
struct NttPassArgs {
    uint32_t count, reg_rounds;
    NttJob job[kJobsPerLaunch];
};

// ---- the in-tile DFT with each wave's 256 elements in registers ----------
// A wave holds 256 consecutive tile elements e = (group << deg) | i, four per
// lane.  Eight "slots" address them: lane bits 0..5 and register bits
// rho0, rho1 (register r = rho0 + 2 rho1).  Radix-2 DIF rounds run from the
// top bit of i down; a round whose bit sits in a register slot pairs
// registers of one lane, a round whose bit sits in lane slot L first swaps
// that lane bit with a register slot whose bit is done: lane distance 32 and
// 16 by v_permlane32_swap / v_permlane16_swap, 8 / 4 / 2 / 1 by DPP
// exchanges.  No LDS traffic and no barrier between rounds.
//
// Initial slots: lane bit b holds e-bit b for b <= deg - 3 (the low bits of
// i, so consecutive lanes read consecutive LDS elements) and e-bit b + 2
// above (the group bits); rho0 = e-bit deg - 2, rho1 = e-bit deg - 1.  Round
// r >= 2 handles i-bit deg - 1 - r, found in lane slot deg - 1 - r, and
// swaps it with rho0 (r even) or rho1 (r odd).
struct SlotMap {
    int e[8];  // e-bit held by slot: 0..5 lane bits, 6 rho0, 7 rho1
};
__host__ __device__ constexpr SlotMap slots_initial(int deg) {
    SlotMap s{};
    for (int b = 0; b < 6; b++) s.e[b] = b <= deg - 3 ? b : b + 2;
    s.e[6] = deg - 2;
    s.e[7] = deg - 1;
    return s;
}
__host__ __device__ constexpr SlotMap slots_final(int deg) {
    SlotMap s = slots_initial(deg);
    for (int r = 2; r < deg; r++) {
        const int b = deg - 1 - r, k = 6 + (r & 1);
        const int t = s.e[b];
        s.e[b] = s.e[k];
        s.e[k] = t;
    }
    return s;
}
template <int DEG, bool FINAL>
__device__ __forceinline__ uint32_t slot_elem(uint32_t lane, uint32_t reg) {
    constexpr SlotMap m = FINAL ? slots_final(DEG) : slots_initial(DEG);
    uint32_t e = 0;
#pragma unroll
    for (int b = 0; b < 6; b++) e |= ((lane >> b) & 1u) << m.e[b];
    e |= (reg & 1u) << m.e[6];
    e |= (reg >> 1) << m.e[7];
    return e;
}

// v from lane (lane ^ 2^L), L <= 3, by DPP within rows of 16 lanes:
// xor 1 / xor 2 as quad permutations, xor 4 = half-row mirror of the
// reversed quads, xor 8 = half-row mirror of the row mirror
template <int L>
__device__ __forceinline__ uint32_t lanes_xor(uint32_t v) {
    const int x = (int)v;
    if constexpr (L == 0) {
        return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);
    } else if constexpr (L == 1) {
        return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);
    } else if constexpr (L == 2) {
        return (uint32_t)__builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp(x, 0x1B, 0xF, 0xF, false), 0x141, 0xF,
                                                  0xF, false);
    } else {
        return (uint32_t)__builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp(x, 0x140, 0xF, 0xF, false), 0x141, 0xF,
                                                  0xF, false);
    }
}

// Swaps lane bit L with the register bit that tells ra (bit 0) from rb
// (bit 1): afterwards ra / rb of a lane with lane bit L = 0 hold the old
// ra of this lane / of its partner, and with lane bit L = 1 the old rb of
// the partner / of this lane.
template <int L>
__device__ __forceinline__ void swap_lane_reg(Fr &ra, Fr &rb, uint32_t lane) {
    if constexpr (L == 5) {
#pragma unroll
        for (int d = 0; d < 8; d++) {
            const auto r = __builtin_amdgcn_permlane32_swap(ra.v[d], rb.v[d], false, false);
            ra.v[d] = r[0];
            rb.v[d] = r[1];
        }
    } else if constexpr (L == 4) {
#pragma unroll
        for (int d = 0; d < 8; d++) {
            const auto r = __builtin_amdgcn_permlane16_swap(ra.v[d], rb.v[d], false, false);
            ra.v[d] = r[0];
            rb.v[d] = r[1];
        }
    } else {
        const bool up = (lane >> L) & 1u;
#pragma unroll
        for (int d = 0; d < 8; d++) {
            const uint32_t recv = lanes_xor<L>(up ? ra.v[d] : rb.v[d]);
            // two selects, not a store through a selected address (which
            // would move ra / rb to scratch memory)
            ra.v[d] = up ? recv : ra.v[d];
            rb.v[d] = up ? rb.v[d] : recv;
        }
    }
}

// Register rounds keep values below 2r (r < 2^254, so 4r < 2^256): a sum is
// brought below 2r by one conditional subtraction of 2r; the difference
// that feeds a twiddle product is a - c + 2r (below 4r, no condition), and
// the 29-bit product of it and a twiddle in the form f29_from32 gives (below
// 32r) is below 4r * 32r / 2^261 + r < 2r, so its final reduction is left
// out.  wave_dft_regs reduces every value below r once, after the rounds.
__device__ __forceinline__ uint32_t fr_2r_limb(int i) {
    return (FrParams::M[i] << 1) | (i ? FrParams::M[i - 1] >> 31 : 0u);
}
// a + b, then - 2r unless that borrows (a, b < 2r -> < 2r)
__device__ __forceinline__ Fr fr_add_lazy(const Fr &a, const Fr &b) {
    Fr t, u;
    uint32_t c = 0, br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) t.v[i] = add_cc(a.v[i], b.v[i], c);
#pragma unroll
    for (int i = 0; i < 8; i++) u.v[i] = sub_bb(t.v[i], fr_2r_limb(i), br);
#pragma unroll
    for (int i = 0; i < 8; i++) t.v[i] = br ? t.v[i] : u.v[i];
    return t;
}
// a - b, plus 2r when that borrows (a, b < 2r -> < 2r)
__device__ __forceinline__ Fr fr_sub_lazy(const Fr &a, const Fr &b) {
    Fr t;
    uint32_t br = 0, c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) t.v[i] = sub_bb(a.v[i], b.v[i], br);
    const uint32_t mask = 0u - br;
#pragma unroll
    for (int i = 0; i < 8; i++) t.v[i] = add_cc(t.v[i], fr_2r_limb(i) & mask, c);
    return t;
}
// a + 2r - b (a, b < 2r -> in (0, 4r)), a product operand only
__device__ __forceinline__ Fr fr_sub_2r(const Fr &a, const Fr &b) {
    Fr t;
    uint32_t c = 0, br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) t.v[i] = add_cc(a.v[i], fr_2r_limb(i), c);
#pragma unroll
    for (int i = 0; i < 8; i++) t.v[i] = sub_bb(t.v[i], b.v[i], br);
    return t;
}

// w: a twiddle in the 29-bit form f29_from32 gives (register rounds), unused
// without MUL.  a, c < 2r in and out.
template <bool MUL, class W>
__device__ __forceinline__ void butterfly(Fr &a, Fr &c, const W &w) {
    const Fr s = fr_add_lazy(a, c);
    if constexpr (MUL) {
        const F29 t = f29_mul<Fr29Mod>(f29_repack(fr_sub_2r(a, c).v), w);
        f29_unpack(t, c.v);
    } else {
        c = fr_sub_lazy(a, c);
    }
    a = s;
}

// Rounds RND..DEG-1 (RND >= 2), one template instance per round so that
// every register index and lane distance is a compile-time constant.
template <int DEG, int RND>
__device__ __forceinline__ void dft_rounds(Fr &r0, Fr &r1, Fr &r2, Fr &r3, const F29 *tw, uint32_t lane) {
    if constexpr (RND < DEG) {
        constexpr int b = DEG - 1 - RND;  // i-bit of this round, in lane slot b
        if constexpr ((RND & 1) == 0) {   // swap with rho0: pairs (0,1), (2,3)
            swap_lane_reg<b>(r0, r1, lane);
            swap_lane_reg<b>(r2, r3, lane);
            if constexpr (b > 0) {
                const F29 w = tw[(lane & ((1u << b) - 1)) << RND];
                butterfly<true>(r0, r1, w);
                butterfly<true>(r2, r3, w);
            } else {  // last round: twiddle 1
                butterfly<false>(r0, r1, r0);
                butterfly<false>(r2, r3, r0);
            }
        } else {  // swap with rho1: pairs (0,2), (1,3)
            swap_lane_reg<b>(r0, r2, lane);
            swap_lane_reg<b>(r1, r3, lane);
            if constexpr (b > 0) {
                const F29 w = tw[(lane & ((1u << b) - 1)) << RND];
                butterfly<true>(r0, r2, w);
                butterfly<true>(r1, r3, w);
            } else {
                butterfly<false>(r0, r2, r0);
                butterfly<false>(r1, r3, r0);
            }
        }
        dft_rounds<DEG, RND + 1>(r0, r1, r2, r3, tw, lane);
    }
}

// The DIF of every group of 2^DEG in the wave's 256 elements (s: the wave's
// LDS region, e-indexed; tw: w_D^j for j < D/2 in the 29-bit form of
// f29_from32, each twiddle product a 29-bit one).  Results go back to s in
// the layout the radix-2 LDS rounds leave (U[bitrev(i)] at i).  quarter:
// (wave-uniform) every element with i >= D/4 is zero, as in the first pass
// of a coset extension by 4 -- those are r1..r3 (rho != 0), so rounds 0 and 1
// reduce to three products and no additions.
template <int DEG>
__device__ __forceinline__ void wave_dft_regs(Fr *s, const F29 *tw, uint32_t lane, bool quarter) {
    static_assert(DEG >= 2 && DEG <= 8, "register DFT covers 2 <= deg <= 8");
    const uint32_t lo = lane & ((1u << (DEG - 2)) - 1);
    Fr r0 = s[slot_elem<DEG, false>(lane, 0)], r1, r2, r3;
    if (quarter) {
        // round 0: (r0, 0) -> (r0, r0 w); round 1: (x, 0) -> (x, x w')
        r2 = fr_mul_w29(r0, tw[lo]);
        if constexpr (DEG == 2) {
            r1 = r0;
            r3 = r2;
        } else {
            const F29 w = tw[lo << 1];
            r1 = fr_mul_w29(r0, w);
            r3 = fr_mul_w29(r2, w);
        }
    } else {
        r1 = s[slot_elem<DEG, false>(lane, 1)];
        r2 = s[slot_elem<DEG, false>(lane, 2)];
        r3 = s[slot_elem<DEG, false>(lane, 3)];
        // round 0: i-bit DEG-1 in rho1; pairs (0,2) and (1,3)
        butterfly<true>(r0, r2, tw[lo]);
        butterfly<true>(r1, r3, tw[lo + (1u << (DEG - 2))]);
        // round 1: i-bit DEG-2 in rho0; pairs (0,1) and (2,3)
        if constexpr (DEG == 2) {
            butterfly<false>(r0, r1, r0);
            butterfly<false>(r2, r3, r0);
        } else {
            const F29 w = tw[lo << 1];
            butterfly<true>(r0, r1, w);
            butterfly<true>(r2, r3, w);
        }
    }
    dft_rounds<DEG, 2>(r0, r1, r2, r3, tw, lane);
    // the rounds leave values below 2r: canonical again for the LDS
    s[slot_elem<DEG, true>(lane, 0)] = fe_reduce_once<FrParams>(r0);
    s[slot_elem<DEG, true>(lane, 1)] = fe_reduce_once<FrParams>(r1);
    s[slot_elem<DEG, true>(lane, 2)] = fe_reduce_once<FrParams>(r2);
    s[slot_elem<DEG, true>(lane, 3)] = fe_reduce_once<FrParams>(r3);
}

// Register rounds for 1024-element tiles of degree 2..8 (the radix-2 LDS
// rounds below stay for smaller tiles and degree 1).  ZK_NTT_ROUNDS=lds
// selects the LDS rounds everywhere (A/B measurement).
__host__ __device__ inline bool use_register_rounds(uint32_t tile, uint32_t deg) {
    return tile == kTile && deg >= 2 && deg <= 8;
}

__global__ __launch_bounds__(kThreads) void k_ntt_pass(const NttPassArgs a) {
    extern __shared__ Fr lds[];
    // the job of this block (uniform: binary search in the kernel arguments)
    uint32_t lo = 0, hi = a.count;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (a.job[mid].block0 <= blockIdx.x) lo = mid; else hi = mid;
    }
    const Fr *__restrict__ x = a.job[lo].x;
    Fr *__restrict__ y = a.job[lo].y;
    const Fr *__restrict__ omegas = a.job[lo].omegas;
    const Fr *__restrict__ n_inv = a.job[lo].n_inv;
    const uint32_t log_n = a.job[lo].log_n, log_p = a.job[lo].log_p, deg = a.job[lo].deg, log_g = a.job[lo].log_g;
    const uint32_t nz = a.job[lo].nz, cn = a.job[lo].cn;
    const Fr *__restrict__ coset = a.job[lo].coset;
    const uint32_t blk = blockIdx.x - a.job[lo].block0;
    const uint32_t tid = threadIdx.x, bs = blockDim.x;
    const uint32_t D = 1u << deg;
    const uint32_t G = 1u << log_g;
    const uint32_t tile = D << log_g;
    const uint32_t log_t = log_n - deg;
    const uint32_t p = 1u << log_p;
    const uint32_t index0 = blk << log_g;
    Fr *s = lds;
    Fr *tw = lds + tile;                              // radix-2 LDS rounds: 32-bit twiddles
    F29 *tw29 = reinterpret_cast<F29 *>(lds + tile);  // register rounds: 29-bit twiddles
    const bool regs = a.reg_rounds && use_register_rounds(tile, deg);

    for (uint32_t j = tid; j < D / 2; j += bs) {
        const Fr w = fe_load<FrParams>(&omegas[j << log_t]);
        if (regs)
            tw29[j] = f29_from32(w);
        else
            tw[j] = w;
    }

    const uint32_t base_shift = log_t - log_p;
    // the coset scales (at most two distinct values: zeta, zeta^2) in the
    // 29-bit product's form, converted once
    F29 cs1{}, cs2{};
    if (cn > 1) {
        cs1 = f29_from32(fe_load<FrParams>(&coset[0]));
        cs2 = cn > 2 ? f29_from32(fe_load<FrParams>(&coset[1])) : cs1;
    }
    for (uint32_t e = tid; e < tile; e += bs) {
        uint32_t g = e & (G - 1), i = e >> log_g;
        uint32_t index = index0 + g;
        const size_t pos = (size_t)index + ((size_t)i << log_t);
        Fr v;
        if (nz && pos >= nz) {
            v = fe_zero<FrParams>();
        } else {
            v = fe_load<FrParams>(&x[pos]);
            if (cn > 1) {
                const uint32_t r = (uint32_t)(pos % cn);
                if (r == 1 || r == 2) {
                    F29 w;  // selected limb by limb: no indexed local array
#pragma unroll
                    for (int q = 0; q < 9; q++) w.l[q] = r == 1 ? cs1.l[q] : cs2.l[q];
                    v = fr_mul_w29(v, w);
                } else if (r) {
                    v = v * fe_load<FrParams>(&coset[r - 1]);
                }
            }
        }
        if (log_p) {
            uint32_t k = index & (p - 1);
            uint32_t ex = (k * i) << base_shift;
            if (ex) v = fr_mul_w29(v, f29_from32(fe_load<FrParams>(&omegas[ex])));
        }
        s[(g << deg) + i] = v;
    }
    __syncthreads();

    if (regs) {
        // bs == kThreads: four waves, each owning 256 elements of the tile
        Fr *ws = s + ((tid >> 6) << 8);
        const uint32_t lane = tid & 63;
        // first pass of a zero-padded input with nz <= n / 4: i >= D/4 reads
        // positions index + i * t >= (D/4) * t = n / 4 >= nz, all zero
        const bool quarter = nz && log_p == 0 && ((size_t)nz << 2) <= ((size_t)1 << log_n);
        switch (deg) {
            case 2: wave_dft_regs<2>(ws, tw29, lane, quarter); break;
            case 3: wave_dft_regs<3>(ws, tw29, lane, quarter); break;
            case 4: wave_dft_regs<4>(ws, tw29, lane, quarter); break;
            case 5: wave_dft_regs<5>(ws, tw29, lane, quarter); break;
            case 6: wave_dft_regs<6>(ws, tw29, lane, quarter); break;
            case 7: wave_dft_regs<7>(ws, tw29, lane, quarter); break;
            default: wave_dft_regs<8>(ws, tw29, lane, quarter); break;
        }
        __syncthreads();
    } else {
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
    }

    const uint32_t lp = log_p < log_g ? log_p : log_g;
    const bool scale = n_inv != nullptr;
    const F29 ninv29 = f29_from32(scale ? fe_load<FrParams>(n_inv) : fe_zero<FrParams>());
    for (uint32_t e = tid; e < tile; e += bs) {
        uint32_t klow = e & ((1u << lp) - 1);
        uint32_t j = (e >> lp) & (D - 1);
        uint32_t hi = e >> (lp + deg);
        uint32_t g = (hi << lp) + klow;
        uint32_t index = index0 + g;
        uint32_t k = index & (p - 1);
        size_t addr = ((size_t)(index - k) << deg) + k + ((size_t)j << log_p);
        Fr v = s[(g << deg) + bitrev32(j, deg)];
        if (scale) v = fr_mul_w29(v, ninv29);
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

// Largest pass degree per transform size; 0 = kMaxDeg.  Set by the
// autotuner (zk_ntt_set_max_deg), read by every plan, possibly from other
// threads: atomic (the plan changes only the speed, never the result).
std::atomic<uint8_t> g_max_deg[31];

// Radix plan: ceil(log_n / max_deg) passes, bits spread evenly (larger first).
int ntt_plan(uint32_t log_n, uint32_t *degs) {
    if (log_n == 0) return 0;
    const uint8_t tuned = log_n < 31 ? g_max_deg[log_n].load(std::memory_order_relaxed) : 0;
    const uint32_t md = tuned ? tuned : kMaxDeg;
    uint32_t passes = (log_n + md - 1) / md;
    uint32_t rem = log_n;
    for (uint32_t i = 0; i < passes; i++) {
        uint32_t left = passes - i;
        degs[i] = (rem + left - 1) / left;
        rem -= degs[i];
    }
    return (int)passes;
}

// One transform of a batch: input in a, the other buffer b.  With src set,
// the first pass reads its input from src instead (nz coefficients, the rest
// zero, scaled by coset as in NttJob) and a only receives later passes.
struct NttSpec {
    Fr *a, *b;
    const Fr *omegas, *n_inv;
    uint32_t log_n;
    const Fr *src = nullptr, *coset = nullptr;
    uint32_t nz = 0, cn = 0;
};

// Runs a batch of NTTs: pass r of every transform that has one goes into the
// same launch(es), kJobsPerLaunch transforms per launch.  passes[m] receives
// transform m's pass count (its result is in b when odd).
// ZK_NTT_ROUNDS=lds: radix-2 rounds through LDS for every tile (A/B runs)
static uint32_t register_rounds_enabled() {
    static const uint32_t on = [] {
        const char *e = getenv("ZK_NTT_ROUNDS");
        return (e && strcmp(e, "lds") == 0) ? 0u : 1u;
    }();
    return on;
}

hipError_t ntt_run_batch(const NttSpec *specs, uint32_t count, hipStream_t stream, int *passes) {
    std::vector<uint32_t> degs((size_t)count * 32);
    std::vector<int> np(count);
    std::vector<uint32_t> logp(count, 0);
    int rounds = 0;
    for (uint32_t m = 0; m < count; m++) {
        np[m] = ntt_plan(specs[m].log_n, &degs[(size_t)m * 32]);
        if (np[m] > rounds) rounds = np[m];
        if (passes) passes[m] = np[m];
    }
    for (int r = 0; r < rounds; r++) {
        NttPassArgs args;
        memset(&args, 0, sizeof(args));
        args.reg_rounds = register_rounds_enabled();
        uint32_t blocks = 0;
        size_t lds = 0;
        auto flush = [&]() -> hipError_t {
            if (!args.count) return hipSuccess;
            k_ntt_pass<<<blocks, kThreads, lds, stream>>>(args);
            ZK_CHECK(hipGetLastError());
            memset(&args, 0, sizeof(args));
            args.reg_rounds = register_rounds_enabled();
            blocks = 0;
            lds = 0;
            return hipSuccess;
        };
        for (uint32_t m = 0; m < count; m++) {
            if (r >= np[m]) continue;
            const NttSpec &sp = specs[m];
            const uint32_t n = 1u << sp.log_n;
            const uint32_t deg = degs[(size_t)m * 32 + r];
            uint32_t tile = n < kTile ? n : kTile;
            uint32_t log_tile = log2_exact(tile);
            uint32_t log_g = log_tile > deg ? log_tile - deg : 0;
            uint32_t tile_elems = (1u << deg) << log_g;
            NttJob &j = args.job[args.count++];
            j.x = (r & 1) ? sp.b : sp.a;
            j.y = (r & 1) ? sp.a : sp.b;
            if (r == 0 && sp.src) {
                j.x = sp.src;
                j.nz = sp.nz;
                j.coset = sp.coset;
                j.cn = sp.cn;
            }
            j.omegas = sp.omegas;
            j.n_inv = r == np[m] - 1 ? sp.n_inv : nullptr;
            j.block0 = blocks;
            j.log_n = (uint8_t)sp.log_n;
            j.log_p = (uint8_t)logp[m];
            j.deg = (uint8_t)deg;
            j.log_g = (uint8_t)log_g;
            blocks += n / tile_elems;
            // the tile, then D/2 twiddles (29-bit form: 36 B each)
            size_t l = (size_t)tile_elems * sizeof(Fr) + (size_t)((1u << deg) / 2) * sizeof(F29);
            if (l > lds) lds = l;
            logp[m] += deg;
            if (args.count == kJobsPerLaunch) ZK_CHECK(flush());
        }
        ZK_CHECK(flush());
    }
    return hipSuccess;
}

// Runs the passes of one NTT, ping-ponging between a and b; returns the
// number of passes (the result is in b when odd).
hipError_t ntt_run(Fr *a, Fr *b, const Fr *omegas, const Fr *n_inv, uint32_t log_n,
                   hipStream_t stream, int *passes_out) {
    NttSpec sp{a, b, omegas, n_inv, log_n};
    int passes = 0;
    ZK_CHECK(ntt_run_batch(&sp, 1, stream, &passes));
    if (passes_out) *passes_out = passes;
    return hipSuccess;
}

hipError_t ntt_batch_from(const void *const *src, void *const *buf, void *const *tmp, const void *omegas,
                          const void *n_inv, int log_n, int count, unsigned char *swapped, hipStream_t stream) {
    if (count <= 0 || log_n < 1 || log_n > 30 || !omegas) return hipErrorInvalidValue;
    std::vector<NttSpec> specs(count);
    for (int m = 0; m < count; m++) {
        if (!src[m] || !buf[m] || !tmp[m] || src[m] == buf[m] || src[m] == tmp[m]) return hipErrorInvalidValue;
        specs[m] = NttSpec{(Fr *)buf[m], (Fr *)tmp[m], (const Fr *)omegas, (const Fr *)n_inv, (uint32_t)log_n};
        specs[m].src = (const Fr *)src[m];
        specs[m].nz = 1u << log_n;
    }
    std::vector<int> passes(count);
    ZK_CHECK(ntt_run_batch(specs.data(), (uint32_t)count, stream, passes.data()));
    for (int m = 0; m < count; m++) swapped[m] = (passes[m] & 1) != 0;
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
    return zk::guard([&]() -> hipError_t {
        (void)pq;
        (void)max_deg;
        if (log_n < 0 || log_n > 30 || !buf || !omegas) return hipErrorInvalidValue;
        if (log_n > 0 && !tmp) return hipErrorInvalidValue;
        int passes = 0;
        ZK_CHECK(ntt_run((Fr *)buf, (Fr *)tmp, (const Fr *)omegas, (const Fr *)n_inv, (uint32_t)log_n,
                         stream, &passes));
        if (swap) *swap = (passes & 1) != 0;
        return hipSuccess;
    });
}

// cuda/bn254.cu:1333-1355.  buf[0] = 1 and buf[1] = w are set by the caller
// (src/cuda/ntt.rs:21-23); fills buf[i] = w^i for all i < 2^log_n in natural
// order.  The reference stores a bit-reversed half table for log_n in {22, 24}
// because its merge kernel indexes it that way; `ntt` here reads the natural
// table for every size, so may_bit_reversed does not change the layout.
hipError_t expand_omega_buffer(void *buf, int log_n, int may_bit_reversed) {
    (void)may_bit_reversed;
    return zk_expand_omega_buffer_on(buf, log_n, 0);
}

// expand_omega_buffer on a given stream (the reference entry uses the legacy
// default stream, which serialises with every other blocking stream)
hipError_t zk_expand_omega_buffer_on(void *buf, int log_n, hipStream_t stream) {
    if (log_n < 0 || log_n > 30 || !buf) return hipErrorInvalidValue;
    uint32_t n = 1u << log_n;
    if (n <= 2) return hipSuccess;
    uint32_t chunk = n >= (1u << 16) ? 64 : 8;
    uint32_t threads = (n + chunk - 1) / chunk;
    k_expand_omega<<<(threads + 63) / 64, 64, 0, stream>>>((Fr *)buf, n, chunk);
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

// Batch of independent NTTs (sizes may differ): pass r of every transform
// goes into one launch (kJobsPerLaunch transforms per launch), so a batch of
// m transforms of at most P passes takes P launches instead of m * P.
// Replaces a loop of `ntt` calls (src/cuda/ntt.rs:61-93 per polynomial).
// items is a host array; swapped[m] (host) is set when result m is in tmp.
// Asynchronous on stream, like `ntt`.
hipError_t zk_ntt_batch(const zk_ntt_item *items, int count, unsigned char *swapped, hipStream_t stream) {
    return zk::guard([&]() -> hipError_t {
        if (count <= 0 || !items || !swapped) return hipErrorInvalidValue;
        std::vector<NttSpec> specs(count);
        for (int m = 0; m < count; m++) {
            const zk_ntt_item &it = items[m];
            if (it.log_n < 0 || it.log_n > 30 || !it.buf || !it.omegas || (it.log_n > 0 && !it.tmp))
                return hipErrorInvalidValue;
            specs[m] = NttSpec{(Fr *)it.buf, (Fr *)it.tmp, (const Fr *)it.omegas, (const Fr *)it.n_inv,
                               (uint32_t)it.log_n};
        }
        std::vector<int> passes(count);
        ZK_CHECK(ntt_run_batch(specs.data(), (uint32_t)count, stream, passes.data()));
        for (int m = 0; m < count; m++) swapped[m] = (passes[m] & 1) != 0;
        return hipSuccess;
    });
}

// The coset extension in one pipeline: for each item, the 2^log_n
// coefficients at src, zero-padded to 2^(log_n + log_ext), scaled by
// coset_powers[(i % coset_powers_n) - 1] where i % coset_powers_n != 0 and
// transformed with omegas (the 2^(log_n + log_ext) w^i table) -- the same
// values as copying src into buf, extended_prepare(buf, coset_powers,
// coset_powers_n, 2^log_n, 2^(log_n + log_ext), 0) and ntt(buf, tmp).  The
// first pass reads only the coefficients and scales them as it loads, so the
// zero padding is never written or read.  The result lands in buf, or in tmp
// when swapped[m] is set; src may not alias buf or tmp.
hipError_t zk_coset_extend_batch(const zk_coset_item *items, int count, const void *coset_powers,
                                 int coset_powers_n, unsigned char *swapped, hipStream_t stream) {
    return zk::guard([&]() -> hipError_t {
        if (count <= 0 || !items || !swapped || coset_powers_n < 0 || (coset_powers_n > 1 && !coset_powers))
            return hipErrorInvalidValue;
        std::vector<NttSpec> specs(count);
        for (int m = 0; m < count; m++) {
            const zk_coset_item &it = items[m];
            if (it.log_n < 0 || it.log_ext < 0 || it.log_n + it.log_ext < 1 || it.log_n + it.log_ext > 30 || !it.src ||
                !it.buf || !it.tmp || !it.omegas || it.src == it.buf || it.src == it.tmp)
                return hipErrorInvalidValue;
            NttSpec &sp = specs[m];
            sp = NttSpec{(Fr *)it.buf, (Fr *)it.tmp, (const Fr *)it.omegas, nullptr, (uint32_t)(it.log_n + it.log_ext)};
            sp.src = (const Fr *)it.src;
            sp.nz = 1u << it.log_n;
            sp.coset = (const Fr *)coset_powers;
            sp.cn = coset_powers_n > 1 ? (uint32_t)coset_powers_n : 0u;
        }
        std::vector<int> passes(count);
        ZK_CHECK(ntt_run_batch(specs.data(), (uint32_t)count, stream, passes.data()));
        for (int m = 0; m < count; m++) swapped[m] = (passes[m] & 1) != 0;
        return hipSuccess;
    });
}

// Autotuner hook: the largest pass degree (radix 2^max_deg) for transforms
// of 2^log_n elements; 0 restores the default (8).
hipError_t zk_ntt_set_max_deg(int log_n, int max_deg) {
    if (log_n < 1 || log_n > 30 || max_deg < 0 || max_deg > (int)kMaxDeg) return hipErrorInvalidValue;
    g_max_deg[log_n].store((uint8_t)max_deg, std::memory_order_relaxed);
    return hipSuccess;
}
int zk_ntt_get_max_deg(int log_n) {
    if (log_n < 1 || log_n > 30) return 0;
    const uint8_t tuned = g_max_deg[log_n].load(std::memory_order_relaxed);
    return tuned ? tuned : (int)kMaxDeg;
}

// Autotuner: times the NTT of 2^log_n elements (buf / tmp device buffers,
// omegas from expand_omega_buffer) for each distinct radix plan with largest
// pass degree 4..8 and records the fastest; ms (optional, 9 floats) receives
// the time per max degree.  The data in buf / tmp is overwritten.
hipError_t zk_ntt_autotune(void *buf, void *tmp, const void *omegas, int log_n, int reps, hipStream_t stream,
                           int *best, float *ms) {
    return zk::guard([&]() -> hipError_t {
        if (log_n < 1 || log_n > 30 || !buf || !tmp || !omegas) return hipErrorInvalidValue;
        if (reps < 1) reps = 1;
        const uint8_t saved = g_max_deg[log_n].load(std::memory_order_relaxed);
        hipEvent_t e0, e1;
        ZK_CHECK(hipEventCreate(&e0));
        hipError_t err = hipEventCreate(&e1);
        if (err != hipSuccess) {
            (void)hipEventDestroy(e0);
            return err;
        }
        if (ms)
            for (int d = 0; d <= (int)kMaxDeg; d++) ms[d] = 0;
        float best_ms = 0;
        int best_d = kMaxDeg;
        uint32_t prev_passes = 0;
        for (int d = (int)kMaxDeg; d >= 4 && err == hipSuccess; d--) {
            g_max_deg[log_n].store((uint8_t)d, std::memory_order_relaxed);
            uint32_t degs[32];
            uint32_t passes = (uint32_t)ntt_plan((uint32_t)log_n, degs);
            if (passes == prev_passes) continue;  // same plan as a larger max degree
            prev_passes = passes;
            int np = 0;
            err = ntt_run((Fr *)buf, (Fr *)tmp, (const Fr *)omegas, nullptr, (uint32_t)log_n, stream, &np);
            if (err == hipSuccess) err = hipEventRecord(e0, stream);
            for (int r = 0; r < reps && err == hipSuccess; r++)
                err = ntt_run((Fr *)buf, (Fr *)tmp, (const Fr *)omegas, nullptr, (uint32_t)log_n, stream, &np);
            if (err == hipSuccess) err = hipEventRecord(e1, stream);
            if (err == hipSuccess) err = hipEventSynchronize(e1);
            float t = 0;
            if (err == hipSuccess) err = hipEventElapsedTime(&t, e0, e1);
            if (err != hipSuccess) break;
            t /= reps;
            if (ms) ms[d] = t;
            if (best_ms == 0 || t < best_ms) {
                best_ms = t;
                best_d = d;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        g_max_deg[log_n].store(err == hipSuccess ? (uint8_t)best_d : saved, std::memory_order_relaxed);
        if (err == hipSuccess && best) *best = best_d;
        return err;
    });
}

}  // extern "C"
