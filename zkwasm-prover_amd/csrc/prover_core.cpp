// Prover core: the device arithmetic of one halo2 proof of the PLONKish
// benchmark circuit, in the phase order of the reference's
// _create_proof_from_advices (src/lib.rs:205-1194) and
// evaluate_h_gates_and_vanishing_construct (src/eval_h.rs:410-1196):
//   1. advice commitments in the Lagrange basis, then iNTT to coefficients
//      (batch_msm_and_intt_ext, src/cuda/msm.rs:73-237)
//   2. lookup multiplicities m(X): prepare_table_lookup + calc_m per input,
//      commitments, iNTT (src/lib.rs:404-530)
//   3. grand products after beta, gamma: logup z (logup_sum_input_inv +
//      eval_logup_z_compose, src/lib.rs:590-745), permutation z per chunk of
//      columns (src/permutation.rs:16-150 on the device), shuffle z
//      (src/shuffle.rs:142-209); commitments and iNTT of every z
//   4. h(X) on the extended coset (factor 4) as a Horner fold in y: gate,
//      permutation (p1, p2, one product term per chunk), logup, shuffle
//      (src/eval_h.rs:639-1196)
//   5. vanishing: h / Z_H, extended iNTT, coset un-distribution, one
//      commitment per quotient piece, h_pieces by Horner in x^n
//      (src/eval_h.rs:464-583)
// Everything runs on the caller's stream with the library's own entry points;
// commitments go through the batched MSM (one pipeline pass per group of
// polynomials) or, for comparison, one MSM call per polynomial.
//
// The circuit: columns a, b, c, d with a*b + c - d = 0 under a selector q on
// the usable rows; `lookups` advice columns range-checked against a fixed
// table column (values i mod 2^table_log); `perm_cols` advice columns joined
// by random copy cycles; `shuffles` pairs of advice columns, the second a
// row permutation of the first on 2^shuffle_log rows.  Rows from
// u = n - (blinding + 1) on are random blinding rows.  Challenges are
// Fiat-Shamir: squeezed from a Blake2b transcript that absorbs every
// commitment and evaluation in the reference's order (theta after the advice,
// beta and gamma after m, y after the grand products, x after the quotient
// pieces, then the opening's).  Because the witness satisfies every
// constraint, h(X) is a polynomial of degree < 3n: the coefficients
// [3n, 4n) of the extended iNTT must vanish, which checks the whole
// composition at any size.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <stdlib.h>

#include <algorithm>
#include <chrono>
#include <utility>
#include <vector>

#include "guard.h"
#include "host_ec.h"
#include "host_fr.h"
#include "zkinternal.h"
#include "zkwasm_prover_kernel.h"

namespace {

#define PCHECK(x)                             \
    do {                                      \
        hipError_t e_ = (x);                  \
        if (e_ != hipSuccess) return e_;      \
    } while (0)

constexpr size_t FR = 32;
constexpr int kDegree = 4;             // max constraint degree of the circuit
constexpr int kChunk = kDegree - 2;    // permutation columns per set
constexpr int kQuot = kDegree - 1;     // quotient pieces
constexpr int kExtBits = 2;            // extended domain = 4n
constexpr int kMsmGroup = 16;          // polynomials per batched-MSM call
constexpr int kExtGroup = 8;           // coset extensions per batched NTT

// canonical constants (little-endian u64 limbs)
// w28 = 7^((r-1)/2^28), the 2^28-th root of unity halo2 derives its domains from
constexpr uint64_t kW28[4] = {0xd34f1ed960c37c9cull, 0x3215cf6dd39329c8ull, 0x98865ea93dd31f74ull,
                              0x03ddb9f5166d18b7ull};
// zeta = 7^((r-1)/3), a primitive cube root of unity (coset generator)
constexpr uint64_t kZeta[4] = {0x8b17ea66b99c90ddull, 0x5bfc41088d8daaa7ull, 0xb3c4d79d41a91758ull, 0ull};
// delta = 7^(2^28), generator of the permutation column cosets
constexpr uint64_t kDelta[4] = {0x870e56bbe533e9a2ull, 0x5b5f898e5e963f25ull, 0x64ec26aad4c86e71ull,
                                0x09226b6e22c6f0caull};

enum FrOp { OP_ADD = 0, OP_MUL = 1, OP_SUB = 2, OP_INV = 3, OP_SQR = 4, OP_NEG = 5, OP_TO_MONT = 6 };
// field_op's operations
enum FieldOp { FO_ADD = 0, FO_MUL = 1, FO_LEFT = 2, FO_SUB = 3 };

uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// bump allocation inside one device allocation; a dry pass sizes it
struct Arena {
    uint8_t *base = nullptr;
    size_t off = 0;
    uint8_t *take(size_t bytes) {
        size_t a = (off + 255) & ~(size_t)255;
        off = a + (bytes ? bytes : 1);
        return base ? base + a : nullptr;
    }
    void *fr(size_t count) { return take(count * FR); }
};

struct Shape;
int ptr_slots(const Shape &s);
constexpr int kPtrSlots = 1024;  // device pointer area of a job (8 KiB)
constexpr int kMaxGateCols = 32;  // coefficient gates zk::gate_eval_h takes

struct Shape {
    int k, ek, A, F, I, L, P, PA, S, Sp, FG, tlog, slog, blind;
    uint32_t n, N, u, tsize, R;
    // advice column roles
    int col_gate, col_lookup, col_perm, col_shuffle, col_fixed_gate;
};

// The circuit's columns: fixed 0 is the gate selector q, fixed 1 the lookup
// table t, fixed 2.. the coefficients f_j of the gates q (f_j a - e_j) = 0;
// the permutation spans the instance columns, then perm_cols - instance
// advice columns.  Every field must be honoured: shapes this core cannot
// build are rejected.
bool make_shape(const zkp_circuit *c, Shape &s) {
    if (!c || c->k < 4 || c->k > 24 || c->advice < 0 || c->lookups < 0 || c->perm_cols < 0 || c->shuffles < 0 ||
        c->blinding < 1 || c->table_log < 1 || c->shuffle_log < 1 || c->fixed < 2 || c->fixed > 64 ||
        c->instance < 0 || c->instance > c->perm_cols)
        return false;
    s.k = c->k;
    s.ek = c->k + kExtBits;
    s.n = 1u << c->k;
    s.N = 1u << s.ek;
    s.blind = c->blinding;
    if ((uint32_t)c->blinding + 2 >= s.n) return false;
    s.u = s.n - (uint32_t)(c->blinding + 1);
    s.L = c->lookups;
    s.P = c->perm_cols;
    s.I = c->instance;
    s.PA = s.P - s.I;
    s.S = c->shuffles;
    s.Sp = (s.P + kChunk - 1) / kChunk;
    s.F = c->fixed;
    s.FG = s.F - 2;
    s.tlog = c->table_log;
    s.slog = c->shuffle_log;
    s.tsize = (s.tlog < 31 && (1u << s.tlog) < s.u) ? (1u << s.tlog) : s.u;
    s.R = (s.slog < 31 && (1u << s.slog) < s.u) ? (1u << s.slog) : s.u;
    s.col_gate = 0;
    s.col_lookup = 4;
    s.col_perm = 4 + s.L;
    s.col_shuffle = s.col_perm + s.PA;
    s.col_fixed_gate = s.col_shuffle + 2 * s.S;
    s.A = c->advice;
    if (s.A < s.col_fixed_gate + s.FG) return false;
    if (ptr_slots(s) > kPtrSlots) return false;  // device pointer area of a job
    if (s.FG > kMaxGateCols) return false;        // the gate kernel's argument block
    return true;
}

// constants area (Fr slots)
enum Const {
    C_ONE, C_MONE, C_ZERO, C_NINV, C_NNINV, C_ZETA, C_ZETA2, C_DELTA, C_WN, C_WNINV, C_WE, C_WEINV,
    C_FWD0, C_FWD1,  // [zeta, zeta^2]: into the coset
    C_INV0, C_INV1,  // [zeta^2, zeta]: out of the coset
    C_TEV,           // 4 values of 1 / Z_H on the extended coset
    C_TAU = C_TEV + 4,  // the setup's secret (test access only)
    C_DELTAS,
    C_COUNT_FIXED = C_DELTAS  // + P
};

struct Pk {
    Shape sh;
    uint8_t *mem = nullptr;
    size_t bytes = 0;
    void *g_lag, *g, *w_n, *w_n_inv, *w_e, *w_e_inv;
    // fixed-base MSM tables of g_lag and g for window tab_c (0: none), used
    // by the batched commitments
    void *g_lag_tab = nullptr, *g_tab = nullptr;
    int tab_c = 0;
    uint8_t *cst;  // Fr constants
    void *t_lag, *t_ext, *q_ext, *l0_ext, *llast_ext, *lact_ext, *ones_ext, *x_ext;
    void *t_coef, *q_coef;  // fixed columns opened in the multi-open step
    std::vector<void *> fix_lag, fix_coef, fix_ext;  // the gate coefficients f_j
    std::vector<uint8_t> vk;  // commitments of q, t, f_j, sigmas (affine, host)
    zkh::Fr vk_repr{};        // the verifying key's transcript representation (Montgomery)
    std::vector<void *> sig_lag, sig_coef, sig_ext;
    uint32_t *perm_head, *perm_target;
    std::vector<uint32_t *> shuf_perm;
    void **sig_lag_dev;  // device array of P pointers
    void *scratch_n[3];  // setup work buffers (n)
    void *scratch_e;     // setup work buffer (N)
    void *cf(int c) const { return cst + (size_t)c * FR; }
};

// Window of the key's fixed-base MSM tables for n points, 0 without tables.
// Tables pay up to 2^20 points (measured per MSM against the best plain
// window, batches of 16: 2^16 0.16 against 0.23 ms, 2^18 0.49 against 0.58,
// 2^20 1.64 against 1.67); at 2^22 the plain pipeline is faster (its point
// array stays in the GPU's last-level cache, a 4 GB table does not).
// ZKP_MSM_TABLES=0 turns them off, =1 on at every size.
int key_table_window(uint32_t n) {
    const char *e = getenv("ZKP_MSM_TABLES");
    if (e && strcmp(e, "0") == 0) return 0;
    if (n > (1u << 20) && !(e && strcmp(e, "1") == 0)) return 0;
    return zk_msm_table_window((int)n);
}

void pk_layout(Pk &pk, Arena &a) {
    const Shape &s = pk.sh;
    pk.g_lag = a.take((size_t)s.n * 64);
    pk.g = a.take((size_t)s.n * 64);
    pk.tab_c = key_table_window(s.n);
    const size_t tab = pk.tab_c ? zk_msm_table_bytes((int)s.n, pk.tab_c) : 0;
    if (!tab) pk.tab_c = 0;
    if (pk.tab_c) {
        pk.g_lag_tab = a.take(tab);
        pk.g_tab = a.take(tab);
    }
    pk.w_n = a.fr(s.n);
    pk.w_n_inv = a.fr(s.n);
    pk.w_e = a.fr(s.N);
    pk.w_e_inv = a.fr(s.N);
    pk.cst = a.take((size_t)(C_COUNT_FIXED + s.P + 8) * FR);
    pk.t_lag = a.fr(s.n);
    pk.t_ext = a.fr(s.N);
    pk.t_coef = a.fr(s.n);
    pk.q_coef = a.fr(s.n);
    pk.q_ext = a.fr(s.N);
    pk.l0_ext = a.fr(s.N);
    pk.llast_ext = a.fr(s.N);
    pk.lact_ext = a.fr(s.N);
    pk.ones_ext = a.fr(s.N);
    pk.x_ext = a.fr(s.N);
    pk.fix_lag.assign(s.FG, nullptr);
    pk.fix_coef.assign(s.FG, nullptr);
    pk.fix_ext.assign(s.FG, nullptr);
    for (int j = 0; j < s.FG; j++) {
        pk.fix_lag[j] = a.fr(s.n);
        pk.fix_coef[j] = a.fr(s.n);
        pk.fix_ext[j] = a.fr(s.N);
    }
    pk.sig_lag.assign(s.P, nullptr);
    pk.sig_coef.assign(s.P, nullptr);
    pk.sig_ext.assign(s.P, nullptr);
    for (int j = 0; j < s.P; j++) {
        pk.sig_lag[j] = a.fr(s.n);
        pk.sig_coef[j] = a.fr(s.n);
        pk.sig_ext[j] = a.fr(s.N);
    }
    pk.perm_head = (uint32_t *)a.take((size_t)s.P * s.u * 4);
    pk.perm_target = (uint32_t *)a.take((size_t)s.P * s.u * 4);
    pk.shuf_perm.assign(s.S, nullptr);
    for (int j = 0; j < s.S; j++) pk.shuf_perm[j] = (uint32_t *)a.take((size_t)s.R * 4);
    pk.sig_lag_dev = (void **)a.take((size_t)(s.P + 1) * 8);
    for (auto &p : pk.scratch_n) p = a.fr(s.n);
    pk.scratch_e = a.fr(s.N);
}

// canonical limbs -> Montgomery constant slot
hipError_t put_const(Pk &pk, int slot, const uint64_t v[4], hipStream_t st) {
    PCHECK(hipMemcpy(pk.cf(slot), v, FR, hipMemcpyHostToDevice));
    return zk_fr_op(pk.cf(slot), pk.cf(slot), nullptr, OP_TO_MONT, 1, st);
}
hipError_t put_small(void *dst, uint64_t v, hipStream_t st) {
    uint64_t l[4] = {v, 0, 0, 0};
    PCHECK(hipMemcpy(dst, l, FR, hipMemcpyHostToDevice));
    return zk_fr_op(dst, dst, nullptr, OP_TO_MONT, 1, st);
}
// dst = src^(2^times)
hipError_t square_times(void *dst, const void *src, int times, hipStream_t st) {
    if (dst != src) PCHECK(hipMemcpyAsync(dst, src, FR, hipMemcpyDeviceToDevice, st));
    for (int i = 0; i < times; i++) PCHECK(zk_fr_op(dst, dst, nullptr, OP_SQR, 1, st));
    return hipSuccess;
}
// w^i table of 2^log entries for the root w
hipError_t omega_table(void *tab, const void *w, int log, const void *one, hipStream_t st) {
    PCHECK(hipMemcpyAsync(tab, one, FR, hipMemcpyDeviceToDevice, st));
    if (log > 0) PCHECK(hipMemcpyAsync((uint8_t *)tab + FR, w, FR, hipMemcpyDeviceToDevice, st));
    return zk_expand_omega_buffer_on(tab, log, st);
}
// coefficients (n) -> extended coset values (N): zero-pad, distribute
// [zeta, zeta^2], NTT of 2^ek, in one pipeline (the padding is never written)
hipError_t to_coset(const Pk &pk, void *ext, const void *coef, void *tmp_e, void **result, hipStream_t st) {
    const Shape &s = pk.sh;
    zk_coset_item it{coef, ext, tmp_e, pk.w_e, s.k, kExtBits};
    unsigned char sw = 0;
    PCHECK(zk_coset_extend_batch(&it, 1, pk.cf(C_FWD0), 3, &sw, st));
    *result = sw ? tmp_e : ext;
    return hipSuccess;
}

// canonical little-endian bytes of a Montgomery Fp coordinate
void fp_canonical(const uint8_t *mont32, uint8_t *out32) {
    zkh::Fp v;
    memcpy(&v, mont32, 32);
    const zkh::Fp one_raw{{1, 0, 0, 0}};
    v = zkh::fp_mul(v, one_raw);
    memcpy(out32, &v, 32);
}

// The verifying key's transcript representation, absorbed first by every
// proof (vk.hash_into, src/lib.rs:242): Blake2b-512 personalised
// "Halo2-Verify-Key" over the pinned description of the key, read as a
// scalar by from_bytes_wide.  halo2 hashes the Debug text of its pinned
// key; this core hashes its own canonical encoding of the same contents:
// the domain (k, extended k, blinding rows), the column and argument counts,
// the table and shuffle sizes, and every fixed and sigma commitment
// (canonical x, y; zeros for the identity).
void vk_transcript_repr(Pk &pk) {
    const Shape &s = pk.sh;
    std::vector<uint8_t> b;
    const char tag[] = "zkwasm-prover-amd pinned vk v1";
    b.insert(b.end(), tag, tag + sizeof(tag) - 1);
    const int32_t fields[] = {s.k, s.ek, s.blind, s.A, s.F, s.I, s.L, s.P, s.S, s.tlog, s.slog, kDegree};
    for (int32_t f : fields)
        for (int i = 0; i < 4; i++) b.push_back((uint8_t)((uint32_t)f >> (8 * i)));
    for (size_t c = 0; c < pk.vk.size() / 64; c++) {
        const uint8_t *p = pk.vk.data() + c * 64;
        bool id = true;
        for (int i = 0; i < 64; i++) id &= p[i] == 0;
        uint8_t xy[64] = {0};
        if (!id) {
            fp_canonical(p, xy);
            fp_canonical(p + 32, xy + 32);
        }
        b.insert(b.end(), xy, xy + 64);
    }
    static const uint8_t personal[16] = {'H', 'a', 'l', 'o', '2', '-', 'V', 'e', 'r', 'i', 'f', 'y', '-', 'K', 'e', 'y'};
    uint8_t d[64];
    zk_blake2b(d, 64, b.data(), b.size(), personal);
    zk_fr_from_bytes_wide(d, &pk.vk_repr);
}

hipError_t pk_build(Pk &pk, uint64_t seed, hipStream_t st) {
    const Shape &s = pk.sh;
    // constants
    const uint64_t one[4] = {1, 0, 0, 0};
    PCHECK(put_const(pk, C_ONE, one, st));
    PCHECK(hipMemsetAsync(pk.cf(C_ZERO), 0, FR, st));
    PCHECK(zk_fr_op(pk.cf(C_MONE), pk.cf(C_ONE), nullptr, OP_NEG, 1, st));
    PCHECK(put_const(pk, C_ZETA, kZeta, st));
    PCHECK(zk_fr_op(pk.cf(C_ZETA2), pk.cf(C_ZETA), nullptr, OP_SQR, 1, st));
    PCHECK(put_const(pk, C_DELTA, kDelta, st));
    PCHECK(put_const(pk, C_WE, kW28, st));  // temporarily w28
    PCHECK(square_times(pk.cf(C_WN), pk.cf(C_WE), 28 - s.k, st));
    PCHECK(square_times(pk.cf(C_WE), pk.cf(C_WE), 28 - s.ek, st));
    PCHECK(zk_fr_op(pk.cf(C_WNINV), pk.cf(C_WN), nullptr, OP_INV, 1, st));
    PCHECK(zk_fr_op(pk.cf(C_WEINV), pk.cf(C_WE), nullptr, OP_INV, 1, st));
    PCHECK(put_small(pk.cf(C_NINV), s.n, st));
    PCHECK(zk_fr_op(pk.cf(C_NINV), pk.cf(C_NINV), nullptr, OP_INV, 1, st));
    PCHECK(put_small(pk.cf(C_NNINV), s.N, st));
    PCHECK(zk_fr_op(pk.cf(C_NNINV), pk.cf(C_NNINV), nullptr, OP_INV, 1, st));
    PCHECK(hipMemcpyAsync(pk.cf(C_FWD0), pk.cf(C_ZETA), FR, hipMemcpyDeviceToDevice, st));
    PCHECK(hipMemcpyAsync(pk.cf(C_FWD1), pk.cf(C_ZETA2), FR, hipMemcpyDeviceToDevice, st));
    PCHECK(hipMemcpyAsync(pk.cf(C_INV0), pk.cf(C_ZETA2), FR, hipMemcpyDeviceToDevice, st));
    PCHECK(hipMemcpyAsync(pk.cf(C_INV1), pk.cf(C_ZETA), FR, hipMemcpyDeviceToDevice, st));
    // deltas[j] = delta^j
    PCHECK(hipMemcpyAsync(pk.cf(C_DELTAS), pk.cf(C_ONE), FR, hipMemcpyDeviceToDevice, st));
    for (int j = 1; j < s.P; j++)
        PCHECK(zk_fr_op(pk.cf(C_DELTAS + j), pk.cf(C_DELTAS + j - 1), pk.cf(C_DELTA), OP_MUL, 1, st));
    // twiddle tables
    PCHECK(omega_table(pk.w_n, pk.cf(C_WN), s.k, pk.cf(C_ONE), st));
    PCHECK(omega_table(pk.w_n_inv, pk.cf(C_WNINV), s.k, pk.cf(C_ONE), st));
    PCHECK(omega_table(pk.w_e, pk.cf(C_WE), s.ek, pk.cf(C_ONE), st));
    PCHECK(omega_table(pk.w_e_inv, pk.cf(C_WEINV), s.ek, pk.cf(C_ONE), st));
    // t_evaluations: 1 / (X^n - 1) at X = zeta w_e^i depends on i mod 4:
    // zeta^n * (w_e^n)^j - 1, j < 4
    {
        void *zn = pk.scratch_n[0];
        PCHECK(square_times(zn, pk.cf(C_ZETA), s.k, st));  // zeta^(2^k)
        void *wn4 = (uint8_t *)pk.scratch_n[0] + FR;
        PCHECK(square_times(wn4, pk.cf(C_WE), s.k, st));  // w_e^n: a 4th root of unity
        void *acc = (uint8_t *)pk.scratch_n[0] + 2 * FR;
        PCHECK(hipMemcpyAsync(acc, zn, FR, hipMemcpyDeviceToDevice, st));
        for (int j = 0; j < 4; j++) {
            void *t = pk.cf(C_TEV + j);
            PCHECK(zk_fr_op(t, acc, pk.cf(C_ONE), OP_SUB, 1, st));
            PCHECK(zk_fr_op(t, t, nullptr, OP_INV, 1, st));
            PCHECK(zk_fr_op(acc, acc, wn4, OP_MUL, 1, st));
        }
    }
    // KZG setup for a seeded secret tau: g[i] = tau^i G, g_lag[i] = L_i(tau) G
    // with L_i(tau) = w^i (tau^n - 1) / (n (tau - w^i))
    {
        void *tau = pk.cf(C_TAU), *sc = pk.scratch_n[0], *tmp = pk.scratch_n[1];
        PCHECK(zk_rand_fr(tau, 1, seed ^ 0x7a7a7a7aull, st));
        PCHECK(omega_table(sc, tau, s.k, pk.cf(C_ONE), st));  // tau^i
        PCHECK(zk_g1_mul_scalars(pk.g, sc, (int)s.n, st));
        void *c = (uint8_t *)pk.scratch_n[2];  // (tau^n - 1) / n
        PCHECK(square_times(c, tau, s.k, st));
        PCHECK(zk_fr_op(c, c, pk.cf(C_ONE), OP_SUB, 1, st));
        PCHECK(zk_fr_op(c, c, pk.cf(C_NINV), OP_MUL, 1, st));
        PCHECK(field_op(sc, nullptr, 0, tau, pk.w_n, 0, nullptr, (int)s.n, 3, st));  // tau - w^i
        PCHECK(zk_batch_invert(sc, tmp, (int)s.n, st));
        PCHECK(field_op(sc, sc, 0, c, pk.w_n, 0, nullptr, (int)s.n, 1, st));
        PCHECK(zk_g1_mul_scalars(pk.g_lag, sc, (int)s.n, st));
        if (pk.tab_c) {
            PCHECK(zk_msm_table_build(pk.g_lag_tab, pk.g_lag, (int)s.n, pk.tab_c, st));
            PCHECK(zk_msm_table_build(pk.g_tab, pk.g, (int)s.n, pk.tab_c, st));
        }
    }
    // fixed columns: table (Lagrange) and its coset; selector q; l0, l_last,
    // l_active; a column of ones
    void *coef = pk.scratch_n[1], *tmp = pk.scratch_n[2], *res = nullptr;
    bool sw = false;
    auto lag_to_coset = [&](const void *lag, void *ext, void *coef_out = nullptr) -> hipError_t {
        PCHECK(hipMemcpyAsync(coef, lag, (size_t)s.n * FR, hipMemcpyDeviceToDevice, st));
        PCHECK(ntt(coef, tmp, nullptr, pk.w_n_inv, pk.cf(C_NINV), s.k, 8, &sw, st));
        if (coef_out)
            PCHECK(hipMemcpyAsync(coef_out, sw ? tmp : coef, (size_t)s.n * FR, hipMemcpyDeviceToDevice, st));
        PCHECK(to_coset(pk, ext, sw ? tmp : coef, pk.scratch_e, &res, st));
        if (res != ext) PCHECK(hipMemcpyAsync(ext, res, (size_t)s.N * FR, hipMemcpyDeviceToDevice, st));
        return hipSuccess;
    };
    PCHECK(zk_rand_fr(pk.t_lag, (int)s.n, seed ^ 0x3333ull, st));
    PCHECK(zk::wit_table_rows(pk.t_lag, s.u, s.tsize, st));
    PCHECK(lag_to_coset(pk.t_lag, pk.t_ext, pk.t_coef));
    void *lag = pk.scratch_n[0];
    PCHECK(zk::wit_indicator(lag, s.n, 0, s.u, st));  // q = l_active on this circuit
    PCHECK(lag_to_coset(lag, pk.q_ext, pk.q_coef));
    PCHECK(hipMemcpyAsync(pk.lact_ext, pk.q_ext, (size_t)s.N * FR, hipMemcpyDeviceToDevice, st));
    PCHECK(zk::wit_indicator(lag, s.n, 0, 1, st));
    PCHECK(lag_to_coset(lag, pk.l0_ext));
    PCHECK(zk::wit_indicator(lag, s.n, s.u, s.u + 1, st));
    PCHECK(lag_to_coset(lag, pk.llast_ext));
    PCHECK(zk::wit_fill_value(pk.ones_ext, pk.cf(C_ONE), s.N, st));
    // the gate coefficients f_j: seeded values on the usable rows, zero on the
    // rest (keygen leaves unassigned fixed cells zero)
    for (int j = 0; j < s.FG; j++) {
        PCHECK(zk_rand_fr(pk.fix_lag[j], (int)s.n, seed ^ (0x6666ull + (uint64_t)j), st));
        PCHECK(zk::wit_zero_rows(pk.fix_lag[j], s.u, s.n, st));
        PCHECK(lag_to_coset(pk.fix_lag[j], pk.fix_ext[j], pk.fix_coef[j]));
    }
    // copy cycles over the usable cells of the permutation columns: a random
    // half of the cells, shuffled, cut into cycles of 2..4
    if (s.P > 0) {
        const uint64_t cells = (uint64_t)s.P * s.u;
        std::vector<uint32_t> order(cells), head(cells), target(cells);
        for (uint64_t i = 0; i < cells; i++) order[i] = head[i] = target[i] = (uint32_t)i;
        uint64_t rs = seed ^ 0x4444ull;
        for (uint64_t i = cells - 1; i > 0; i--) {
            uint64_t j = splitmix(rs) % (i + 1);
            uint32_t t = order[i];
            order[i] = order[j];
            order[j] = t;
        }
        uint64_t used = cells / 2, pos = 0;
        while (pos + 1 < used) {
            uint64_t len = 2 + splitmix(rs) % 3;
            if (pos + len > used) len = used - pos;
            for (uint64_t t = 0; t < len; t++) {
                uint32_t c = order[pos + t];
                head[c] = order[pos];
                target[c] = order[pos + (t + 1) % len];
            }
            pos += len;
        }
        PCHECK(hipMemcpy(pk.perm_head, head.data(), cells * 4, hipMemcpyHostToDevice));
        PCHECK(hipMemcpy(pk.perm_target, target.data(), cells * 4, hipMemcpyHostToDevice));
        std::vector<void *> sp(pk.sig_lag);
        PCHECK(hipMemcpy(pk.sig_lag_dev, sp.data(), (size_t)s.P * 8, hipMemcpyHostToDevice));
        PCHECK(zk::wit_sigma_rows(pk.sig_lag_dev, pk.perm_target, pk.cf(C_DELTAS), pk.w_n, (uint32_t)s.P, s.n, s.u,
                                  st));
        for (int j = 0; j < s.P; j++) {
            PCHECK(hipMemcpyAsync(pk.sig_coef[j], pk.sig_lag[j], (size_t)s.n * FR, hipMemcpyDeviceToDevice, st));
            PCHECK(ntt(pk.sig_coef[j], tmp, nullptr, pk.w_n_inv, pk.cf(C_NINV), s.k, 8, &sw, st));
            if (sw) PCHECK(hipMemcpyAsync(pk.sig_coef[j], tmp, (size_t)s.n * FR, hipMemcpyDeviceToDevice, st));
            PCHECK(to_coset(pk, pk.sig_ext[j], pk.sig_coef[j], pk.scratch_e, &res, st));
            if (res != pk.sig_ext[j])
                PCHECK(hipMemcpyAsync(pk.sig_ext[j], res, (size_t)s.N * FR, hipMemcpyDeviceToDevice, st));
        }
    }
    // the coset points X = zeta w_e^i: coset evaluations of the polynomial X
    PCHECK(hipMemsetAsync(coef, 0, (size_t)s.n * FR, st));
    PCHECK(hipMemcpyAsync((uint8_t *)coef + FR, pk.cf(C_ONE), FR, hipMemcpyDeviceToDevice, st));
    PCHECK(to_coset(pk, pk.x_ext, coef, pk.scratch_e, &res, st));
    if (res != pk.x_ext) PCHECK(hipMemcpyAsync(pk.x_ext, res, (size_t)s.N * FR, hipMemcpyDeviceToDevice, st));
    // shuffle row permutations of the first R rows
    for (int j = 0; j < s.S; j++) {
        std::vector<uint32_t> p(s.R);
        for (uint32_t i = 0; i < s.R; i++) p[i] = i;
        uint64_t rs = seed ^ (0x5555ull + j);
        for (uint32_t i = s.R - 1; i > 0; i--) {
            uint32_t r = (uint32_t)(splitmix(rs) % (i + 1));
            uint32_t t = p[i];
            p[i] = p[r];
            p[r] = t;
        }
        PCHECK(hipMemcpy(pk.shuf_perm[j], p.data(), (size_t)s.R * 4, hipMemcpyHostToDevice));
    }
    // the verifying key: commitments of the fixed columns (q, t, f_j) and of
    // the permutation's sigma columns
    {
        std::vector<const void *> fixed = {pk.q_coef, pk.t_coef};
        for (auto c : pk.fix_coef) fixed.push_back(c);
        for (auto c : pk.sig_coef) fixed.push_back(c);
        PCHECK(hipStreamSynchronize(st));
        pk.vk.assign(fixed.size() * 64, 0);
        PCHECK(zkp_batch_msm(pk.vk.data(), pk.g, fixed.data(), (int)fixed.size(), (int)s.n, 0));
    }
    vk_transcript_repr(pk);
    return hipStreamSynchronize(st);
}

// ---- multi-open query layout -------------------------------------------------
// Rotations of the opening points in halo2's Rotation order:
// last = -(blinding + 1) < cur = 0 < next = 1.
enum Rot { ROT_LAST = 0, ROT_CUR = 1, ROT_NEXT = 2, ROT_COUNT = 3 };
// Polynomials of the opening (coefficient form): instance, advice, fixed
// (q, t, f_j), the vanishing argument's random polynomial, sigma,
// permutation products, lookup multiplicities and sums, shuffle products,
// and h (the quotient folded in x^n).
struct OpenLayout {
    std::vector<const void *> poly;
    std::vector<std::pair<int, int>> evals;    // (poly, rot): src/lib.rs:1025-1078 order, h first
    std::vector<std::pair<int, int>> queries;  // (poly, rot): src/lib.rs:1094-1176 order
};
// GWC commits one quotient per distinct rotation, Shplonk two polynomials
int open_gwc_points(const Shape &s) { return s.Sp > 1 ? 3 : (s.Sp + s.L + s.S > 0 ? 2 : 1); }
int open_eval_count(const Shape &s) {
    const int zp = s.Sp ? 3 * s.Sp - 1 : 0;
    return 1 + s.I + s.A + s.F + 1 + s.P + zp + 3 * s.L + 2 * s.S;
}
constexpr int kOpenCst = 64;

// ---- per-proof state ---------------------------------------------------------
// beta, gamma, y, x; x^n; logup carry; opening challenges v, y', u; opening
// points x w, x w^u; then beta delta^j per permutation column, then the
// gates' y powers and their negatives
enum JConst {
    J_BETA, J_GAMMA, J_Y, J_X, J_XN, J_LASTZ, J_V, J_SY, J_U, J_XW, J_XWU, J_BDELTA,
    J_COUNT_FIXED = J_BDELTA
};

struct Job {
    const Pk *pk;
    uint64_t seed;
    uint8_t *mem = nullptr;
    size_t bytes = 0;
    std::vector<void *> inst_lag, inst_coef, inst_tmp, adv_lag, adv_coef, adv_tmp;
    std::vector<void *> m_lag, m_coef, m_tmp, zl_lag, zl_coef, zl_tmp, zp_lag, zp_coef, zp_tmp, zs_lag, zs_coef,
        zs_tmp;
    void *rnd_lag, *rnd_coef, *rnd_tmp;  // the vanishing argument's random polynomial
    uint32_t *vals;
    void *w_sum, *w_temp, *w_tab;
    std::vector<void *> ext, ext_tmp;  // coset buffers, NTT spares
    void *h, *h_tmp, *l_acc, *r_acc, *aux0, *aux1;
    uint8_t *cst;
    void **ptrs;  // device pointer arrays
    uint8_t *msm_scratch, *msm_out;
    size_t msm_bytes;
    uint32_t *flag;
    void *evals;     // opening evaluations, one Fr per (polynomial, point)
    void *open_cst;  // opening scalars staged from the host
    std::vector<zkh::Fr> last_evals;  // the last proof's opening (host copies)
    zkh::Fr last_open_ch[3] = {};     // v, y', u
    std::vector<uint8_t> proof;       // the last proof's bytes (written points and scalars)
    // host-resident witness (ZKP_HOST_WITNESS): columns the next proof uploads,
    // the upload stream and one event per commitment group
    std::vector<const void *> host_adv, host_inst;
    hipStream_t upload = nullptr;
    std::vector<hipEvent_t> upload_ev;
    bool p4_staged = false;              // the h(X) round's pointer arrays are on the device
    hipGraphExec_t h_graph = nullptr;    // the captured h(X) round (ZKP_GRAPH)
    hipStream_t capture = nullptr;       // the stream it is recorded on
    zk::CommitFn commit_fn = nullptr;  // batched commitments through a hook (the proof server's)
    void *commit_ctx = nullptr;
    void *cf(int c) const { return cst + (size_t)c * FR; }
};

// how many polynomials get an extended coset in phase 4
int ext_count(const Shape &s) { return 4 + 3 * s.L + s.P + s.Sp + 3 * s.S + s.FG; }
// device pointer area: [values | sigmas | zs] (2P + Sp), permutation sets
// (Sp), permutation values and sigma cosets (2P), shuffle inputs and tables
// (2S)
int ptr_slots(const Shape &s) { return 2 * s.P + s.Sp + s.Sp + 2 * s.P + 2 * s.S + 8; }

// scratch of one batched-MSM call: plain points or the key's tables
size_t msm_group_bytes(const Shape &s) {
    std::vector<zk_msm_item> it(kMsmGroup, zk_msm_item{nullptr, nullptr, (int)s.n, 0});
    const size_t plain = zk_msm_batch_scratch_bytes(it.data(), kMsmGroup, 0);
    const int tc = key_table_window(s.n);
    if (!tc) return plain;
    for (auto &x : it) x.table_c = tc;
    const size_t tab = zk_msm_batch_scratch_bytes(it.data(), kMsmGroup, 0);
    return tab > plain ? tab : plain;
}

void job_layout(Job &j, Arena &a) {
    const Shape &s = j.pk->sh;
    auto cols = [&](std::vector<void *> &v, int c, uint32_t len) {
        v.assign(c, nullptr);
        for (auto &p : v) p = a.fr(len);
    };
    cols(j.inst_lag, s.I, s.n);
    cols(j.inst_coef, s.I, s.n);
    cols(j.inst_tmp, s.I, s.n);
    cols(j.adv_lag, s.A, s.n);
    cols(j.adv_coef, s.A, s.n);
    cols(j.adv_tmp, s.A, s.n);
    cols(j.m_lag, s.L, s.n);
    cols(j.m_coef, s.L, s.n);
    cols(j.m_tmp, s.L, s.n);
    cols(j.zl_lag, s.L, s.n);
    cols(j.zl_coef, s.L, s.n);
    cols(j.zl_tmp, s.L, s.n);
    cols(j.zp_lag, s.Sp, s.n);
    cols(j.zp_coef, s.Sp, s.n);
    cols(j.zp_tmp, s.Sp, s.n);
    cols(j.zs_lag, s.S, s.n);
    cols(j.zs_coef, s.S, s.n);
    cols(j.zs_tmp, s.S, s.n);
    j.rnd_lag = a.fr(s.n);
    j.rnd_coef = a.fr(s.n);
    j.rnd_tmp = a.fr(s.n);
    j.vals = (uint32_t *)a.take((size_t)s.u * 4);
    j.w_sum = a.fr(s.n);
    j.w_temp = a.fr(s.n);
    j.w_tab = a.fr(s.n);
    cols(j.ext, ext_count(s), s.N);
    cols(j.ext_tmp, kExtGroup, s.N);
    j.h = a.fr(s.N);
    j.h_tmp = a.fr(s.N);
    j.l_acc = a.fr(s.N);
    j.r_acc = a.fr(s.N);
    j.aux0 = a.fr(s.N);
    j.aux1 = a.fr(s.N);
    j.cst = a.take((size_t)(J_COUNT_FIXED + s.P + 8) * FR);
    j.ptrs = (void **)a.take((size_t)kPtrSlots * 8);
    j.msm_bytes = msm_group_bytes(s);
    j.msm_scratch = a.take(j.msm_bytes);
    j.msm_out = a.take((size_t)kMsmGroup * ZK_MSM_OUT_BYTES);
    j.flag = (uint32_t *)a.take(256);
    j.evals = a.fr((size_t)open_eval_count(s));
    j.open_cst = a.fr(kOpenCst);
}

// Commitment groups (first polynomial, count) of a batch of `count`
// polynomials: kMsmGroup each, except while the inputs still arrive from host
// memory, where the groups ramp up as 4, 8, 16, ...: the first MSMs start
// after 4 columns' upload instead of 16 (≈ 9 instead of 37 ms at k=22), and
// each group's upload overlaps the previous group's MSMs.
std::vector<std::pair<size_t, int>> commit_groups(size_t count, bool uploaded) {
    std::vector<std::pair<size_t, int>> g;
    size_t size = uploaded ? 4 : (size_t)kMsmGroup;
    for (size_t first = 0; first < count; first += size, size = size * 2 < (size_t)kMsmGroup ? size * 2 : kMsmGroup)
        g.emplace_back(first, (int)(count - first < size ? count - first : size));
    return g;
}

// commitments of polys (device Lagrange or coefficient values of n) against
// bases; out: affine points (host, 64 B each)
hipError_t commit(Job &j, const void *bases, const std::vector<const void *> &polys, uint8_t *out, bool batched,
                  hipStream_t st, bool uploaded = false) {
    const Shape &s = j.pk->sh;
    // the batched path commits against the key's table of these bases
    const void *src = bases;
    int tc = 0;
    if (j.pk->tab_c && (bases == j.pk->g_lag || bases == j.pk->g)) {
        src = bases == j.pk->g_lag ? j.pk->g_lag_tab : j.pk->g_tab;
        tc = j.pk->tab_c;
    }
    const std::vector<std::pair<size_t, int>> groups = commit_groups(polys.size(), uploaded);
    for (size_t gi = 0; gi < groups.size(); gi++) {
        const size_t first = groups[gi].first;
        const int cnt = groups[gi].second;
        if (uploaded) PCHECK(hipStreamWaitEvent(st, j.upload_ev[gi], 0));
        if (batched && j.commit_fn) {
            std::vector<zk_msm_item> it(cnt);
            for (int m = 0; m < cnt; m++) it[m] = zk_msm_item{src, polys[first + m], (int)s.n, tc};
            PCHECK(j.commit_fn(j.commit_ctx, it.data(), cnt, out + first * 64, st));
            continue;
        }
        std::vector<uint8_t> jac((size_t)cnt * 96);
        if (batched) {
            std::vector<zk_msm_item> it(cnt);
            for (int m = 0; m < cnt; m++) it[m] = zk_msm_item{src, polys[first + m], (int)s.n, tc};
            PCHECK(zk::msm_batch_host(j.msm_out, it.data(), cnt, j.msm_scratch, j.msm_bytes, st, jac.data()));
        } else {
            for (int m = 0; m < cnt; m++)
                PCHECK(zk_msm((uint8_t *)j.msm_out + (size_t)m * ZK_MSM_OUT_BYTES, bases, polys[first + m],
                              (int)s.n, 0, j.msm_scratch, j.msm_bytes, st));
            for (int m = 0; m < cnt; m++)
                PCHECK(hipMemcpyAsync(&jac[(size_t)m * 96], (uint8_t *)j.msm_out + (size_t)m * ZK_MSM_OUT_BYTES, 96,
                                      hipMemcpyDeviceToHost, st));
            PCHECK(hipStreamSynchronize(st));
        }
        for (int m = 0; m < cnt; m++) {
            const uint8_t *q = &jac[(size_t)m * 96];
            bool id = true;
            for (int b = 64; b < 96; b++) id &= q[b] == 0;
            uint8_t *o = out + (first + m) * 64;
            if (id)
                memset(o, 0, 64);
            else
                memcpy(o, q, 64);
        }
    }
    return hipSuccess;
}

// Uploads the host witness (advice then instance columns) into the job's
// Lagrange buffers on the upload stream, recording one event per commitment
// group; the upload starts after the work already on st (the previous
// proof's reads of these buffers).
hipError_t upload_witness(Job &j, size_t ncols, hipStream_t st) {
    const Shape &s = j.pk->sh;
    if (!j.upload) PCHECK(hipStreamCreateWithFlags(&j.upload, hipStreamNonBlocking));
    const std::vector<std::pair<size_t, int>> cg = commit_groups(ncols, true);
    const size_t groups = cg.size();
    while (j.upload_ev.size() < groups + 1) {
        hipEvent_t e;
        PCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        j.upload_ev.push_back(e);
    }
    PCHECK(hipEventRecord(j.upload_ev[groups], st));
    PCHECK(hipStreamWaitEvent(j.upload, j.upload_ev[groups], 0));
    for (size_t gi = 0; gi < groups; gi++) {
        for (size_t c = cg[gi].first; c < cg[gi].first + (size_t)cg[gi].second; c++) {
            const bool adv = c < (size_t)s.A;
            void *dst = adv ? j.adv_lag[c] : j.inst_lag[c - s.A];
            const void *src = adv ? j.host_adv[c] : j.host_inst[c - s.A];
            PCHECK(hipMemcpyAsync(dst, src, (size_t)s.n * FR, hipMemcpyHostToDevice, j.upload));
        }
        PCHECK(hipEventRecord(j.upload_ev[gi], j.upload));
    }
    return hipSuccess;
}

// Whether an NTT of 2^log_n ends in its spare buffer (an odd pass count).
bool ntt_ends_in_spare(int log_n) {
    uint32_t degs[32];
    return (zk::ntt_plan((uint32_t)log_n, degs) & 1) != 0;
}

// iNTT of each lag column into coef[i] (tmp[i] a spare).  The buffers are
// handed over so that the result lands in coef[i] (with an odd pass count
// coef[i] is the spare), so the buffer of a column never changes between
// proofs: device pointer arrays and captured graphs stay valid.
hipError_t intt_all(Job &j, const std::vector<void *> &lag, const std::vector<void *> &coef,
                    const std::vector<void *> &tmp, hipStream_t st) {
    const Shape &s = j.pk->sh;
    const Pk &pk = *j.pk;
    size_t c = lag.size();
    if (!c) return hipSuccess;
    const bool odd = ntt_ends_in_spare(s.k);
    const std::vector<void *> &buf = odd ? tmp : coef, &spare = odd ? coef : tmp;
    // the first pass reads the Lagrange values in place: no copy into coef
    std::vector<unsigned char> sw(c);
    PCHECK(zk::ntt_batch_from(lag.data(), buf.data(), spare.data(), pk.w_n_inv, pk.cf(C_NINV), s.k, (int)c, sw.data(),
                              st));
    for (size_t i = 0; i < c; i++) {
        const void *res = sw[i] ? spare[i] : buf[i];
        if (res != coef[i]) PCHECK(hipMemcpyAsync(coef[i], res, (size_t)s.n * FR, hipMemcpyDeviceToDevice, st));
    }
    return hipSuccess;
}

// extended cosets of coefficient polys into j.ext[first..] (j.ext_tmp spares),
// the result always in j.ext[i] as in intt_all
hipError_t extend_all(Job &j, const std::vector<const void *> &coefs, int first, hipStream_t st) {
    const Shape &s = j.pk->sh;
    const Pk &pk = *j.pk;
    const bool odd = ntt_ends_in_spare(s.ek);
    for (size_t g = 0; g < coefs.size(); g += kExtGroup) {
        int cnt = (int)(coefs.size() - g < (size_t)kExtGroup ? coefs.size() - g : kExtGroup);
        std::vector<zk_coset_item> it(cnt);
        for (int m = 0; m < cnt; m++) {
            void *dst = j.ext[first + g + m], *spare = j.ext_tmp[m];
            it[m] = zk_coset_item{coefs[g + m], odd ? spare : dst, odd ? dst : spare, pk.w_e, s.k, kExtBits};
        }
        std::vector<unsigned char> sw(cnt);
        PCHECK(zk_coset_extend_batch(it.data(), cnt, pk.cf(C_FWD0), 3, sw.data(), st));
        for (int m = 0; m < cnt; m++) {
            void *res = sw[m] ? it[m].tmp : it[m].buf;
            if (res != j.ext[first + g + m])
                PCHECK(hipMemcpyAsync(j.ext[first + g + m], res, (size_t)s.N * FR, hipMemcpyDeviceToDevice, st));
        }
    }
    return hipSuccess;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

hipError_t witness(Job &j, bool from_host, hipStream_t st) {
    const Shape &s = j.pk->sh;
    const Pk &pk = *j.pk;
    const uint64_t seed = j.seed;
    if (from_host)  // the caller's columns are uploaded with the commitments
        return zk_rand_fr(j.rnd_lag, (int)s.n, seed * 1000003ull + 0x2000, st);
    for (int c = 0; c < s.A; c++) PCHECK(zk_rand_fr(j.adv_lag[c], (int)s.n, seed * 1000003ull + c, st));
    // instance columns: the public values of the usable rows (the caller's
    // `instances` in the reference, padded with zeros to n, src/lib.rs:246-253)
    for (int c = 0; c < s.I; c++) PCHECK(zk_rand_fr(j.inst_lag[c], (int)s.n, seed * 1000003ull + 0x1000 + c, st));
    // the vanishing argument's random polynomial (src/buffer.rs:185-206),
    // drawn as its values on the domain
    PCHECK(zk_rand_fr(j.rnd_lag, (int)s.n, seed * 1000003ull + 0x2000, st));
    // gate: d = a * b + c on the usable rows
    PCHECK(zk::wit_gate_rows(j.adv_lag[3], j.adv_lag[0], j.adv_lag[1], j.adv_lag[2], s.u, st));
    // coefficient gates: e_j = f_j * a on the usable rows
    for (int g = 0; g < s.FG; g++)
        PCHECK(zk::wit_mul_rows(j.adv_lag[s.col_fixed_gate + g], pk.fix_lag[g], j.adv_lag[s.col_gate], s.u, st));
    // lookups: inputs are table values on the usable rows
    for (int l = 0; l < s.L; l++) {
        PCHECK(zk::wit_lookup_values(j.vals, s.u, s.tsize, seed * 7919ull + l, st));
        PCHECK(zk::wit_u32_rows_to_fr(j.adv_lag[s.col_lookup + l], j.vals, s.u, st));
    }
    // copy cycles over the permutation columns (instance first, then advice):
    // every cell of a cycle takes its head's value
    if (s.P > 0) {
        std::vector<void *> cols(s.P);
        for (int c = 0; c < s.P; c++) cols[c] = c < s.I ? j.inst_lag[c] : j.adv_lag[s.col_perm + c - s.I];
        PCHECK(hipMemcpyAsync(j.ptrs, cols.data(), (size_t)s.P * 8, hipMemcpyHostToDevice, st));
        PCHECK(zk::wit_cycle_values(j.ptrs, pk.perm_head, (uint32_t)s.P, s.u, st));
        PCHECK(hipStreamSynchronize(st));  // cols is copied from host memory
    }
    for (int c = 0; c < s.I; c++) PCHECK(zk::wit_zero_rows(j.inst_lag[c], s.u, s.n, st));
    // shuffles: the second column is a row permutation of the first on R rows,
    // both zero on [R, u)
    for (int t = 0; t < s.S; t++) {
        void *in = j.adv_lag[s.col_shuffle + 2 * t], *sh = j.adv_lag[s.col_shuffle + 2 * t + 1];
        PCHECK(zk::wit_gather(sh, in, pk.shuf_perm[t], s.R, st));
        PCHECK(zk::wit_zero_rows(in, s.R, s.u, st));
        PCHECK(zk::wit_zero_rows(sh, s.R, s.u, st));
    }
    return hipStreamSynchronize(st);
}

// coefficients (ascending) of the polynomial of degree < m through
// (pts[k], vals[k]) (halo2's lagrange_interpolate)
std::vector<zkh::Fr> lagrange(const zkh::Fr *pts, const zkh::Fr *vals, int m) {
    using namespace zkh;
    std::vector<Fr> res(m, fr_zero());
    for (int k = 0; k < m; k++) {
        std::vector<Fr> num(1, fr_one());
        Fr den = fr_one();
        for (int l = 0; l < m; l++) {
            if (l == k) continue;
            std::vector<Fr> nx(num.size() + 1, fr_zero());
            for (size_t i = 0; i < num.size(); i++) {
                nx[i + 1] = fr_add(nx[i + 1], num[i]);
                nx[i] = fr_sub(nx[i], fr_mul(pts[l], num[i]));
            }
            num.swap(nx);
            den = fr_mul(den, fr_sub(pts[k], pts[l]));
        }
        const Fr c = fr_mul(vals[k], fr_inv(den));
        for (int i = 0; i < m; i++) res[i] = fr_add(res[i], fr_mul(num[i], c));
    }
    return res;
}

// ---- Fiat-Shamir (halo2's Blake2bWrite, csrc/transcript.cpp) -------------------
// common_* absorb only; write_* also append to the proof: a point as its
// compressed encoding (canonical x little-endian, the parity of canonical y
// in the top bit of the last byte; the identity as 32 zero bytes), a scalar
// as its canonical 32 little-endian bytes.
struct Transcript {
    void *t = zk_transcript_new();
    std::vector<uint8_t> *proof = nullptr;
    ~Transcript() { zk_transcript_free(t); }
    // affine commitments as commit() writes them (Montgomery x, y; zeros for
    // the identity, which halo2 cannot absorb: absorbed here as x = y = 0)
    void points(const uint8_t *aff, int count, bool write) {
        for (int i = 0; i < count; i++) {
            zkh::Fp x, y;
            memcpy(&x, aff + (size_t)i * 64, 32);
            memcpy(&y, aff + (size_t)i * 64 + 32, 32);
            const zkh::Fp one_raw{{1, 0, 0, 0}};
            const bool id = zkh::fp_is_zero(x) && zkh::fp_is_zero(y);
            if (!id) {
                x = zkh::fp_mul(x, one_raw);  // canonical
                y = zkh::fp_mul(y, one_raw);
            }
            zk_transcript_common_point(t, &x, &y);
            if (write && proof) {
                uint8_t c[32];
                memcpy(c, &x, 32);
                if (!id && (y.l[0] & 1)) c[31] |= 0x80;
                proof->insert(proof->end(), c, c + 32);
            }
        }
    }
    void scalars(const zkh::Fr *v, int count, bool write) {
        const zkh::Fr one_raw{{1, 0, 0, 0}};
        for (int i = 0; i < count; i++) {
            const zkh::Fr c = zkh::fr_mul(v[i], one_raw);
            zk_transcript_common_scalar(t, &c);
            if (write && proof) proof->insert(proof->end(), (const uint8_t *)&c, (const uint8_t *)&c + 32);
        }
    }
    zkh::Fr challenge() {
        zkh::Fr c;
        zk_transcript_squeeze(t, &c, nullptr);
        return c;
    }
};

// challenges to the job's constant slots (synchronous: the host values are
// temporaries)
hipError_t put_challenges(Job &j, int first, const zkh::Fr *v, int count, hipStream_t st) {
    PCHECK(hipMemcpyAsync(j.cf(first), v, (size_t)count * FR, hipMemcpyHostToDevice, st));
    return hipStreamSynchronize(st);
}

// The opening's polynomials with their evaluation order (src/lib.rs:1025-1078:
// h first, then instance, advice, fixed, the random polynomial, sigma, the
// permutation products at x, x w and -- all but the last -- x w^-(b+1), each
// lookup's multiplicity and sum, each shuffle product) and query order
// (src/lib.rs:1094-1176 with permutation_product_open, lookup_open and
// shuffle_open of src/multiopen.rs:564-660).  h(x) is computed but not
// written: the verifier derives it from the constraints.
OpenLayout open_layout(const Job &j, const void *hpoly) {
    const Pk &pk = *j.pk;
    const Shape &s = pk.sh;
    OpenLayout o;
    auto add = [&](const void *p) {
        o.poly.push_back(p);
        return (int)o.poly.size() - 1;
    };
    const int h = add(hpoly);
    std::vector<int> inst, adv, fix, sig, zp, m, zl, zs;
    for (auto p : j.inst_coef) inst.push_back(add(p));
    for (auto p : j.adv_coef) adv.push_back(add(p));
    fix.push_back(add(pk.q_coef));
    fix.push_back(add(pk.t_coef));
    for (auto p : pk.fix_coef) fix.push_back(add(p));
    const int rnd = add(j.rnd_coef);
    for (auto p : pk.sig_coef) sig.push_back(add(p));
    for (auto p : j.zp_coef) zp.push_back(add(p));
    for (auto p : j.m_coef) m.push_back(add(p));
    for (auto p : j.zl_coef) zl.push_back(add(p));
    for (auto p : j.zs_coef) zs.push_back(add(p));
    auto E = [&](int p, int r) { o.evals.push_back({p, r}); };
    auto Q = [&](int p, int r) { o.queries.push_back({p, r}); };
    E(h, ROT_CUR);
    for (int p : inst) E(p, ROT_CUR);
    for (int p : adv) E(p, ROT_CUR);
    for (int p : fix) E(p, ROT_CUR);
    E(rnd, ROT_CUR);
    for (int p : sig) E(p, ROT_CUR);
    for (size_t i = 0; i < zp.size(); i++) {
        E(zp[i], ROT_CUR);
        E(zp[i], ROT_NEXT);
        if (i + 1 != zp.size()) E(zp[i], ROT_LAST);
    }
    for (int l = 0; l < s.L; l++) {
        E(m[l], ROT_CUR);
        E(zl[l], ROT_CUR);
        E(zl[l], ROT_NEXT);
    }
    for (int p : zs) {
        E(p, ROT_CUR);
        E(p, ROT_NEXT);
    }
    for (int p : inst) Q(p, ROT_CUR);
    for (int p : adv) Q(p, ROT_CUR);
    for (int p : zp) {
        Q(p, ROT_CUR);
        Q(p, ROT_NEXT);
    }
    for (size_t i = zp.size(); i-- > 1;) Q(zp[i - 1], ROT_LAST);  // all but the last, in reverse
    for (int l = 0; l < s.L; l++) {
        Q(m[l], ROT_CUR);
        Q(zl[l], ROT_CUR);
        Q(zl[l], ROT_NEXT);
    }
    for (int p : zs) {
        Q(p, ROT_CUR);
        Q(p, ROT_NEXT);
    }
    for (int p : fix) Q(p, ROT_CUR);
    for (int p : sig) Q(p, ROT_CUR);
    Q(h, ROT_CUR);
    Q(rnd, ROT_CUR);
    return o;
}

// Multi-open of every query (src/multiopen.rs): evaluations of each
// polynomial at its points (batch_poly_eval, src/eval_poly.rs:18), then GWC
// (queries grouped by rotation, one quotient per group, :16-177) or Shplonk
// (polynomials grouped by their rotation sets, h(X) and L(X), :178-561) with
// the divisions by (X - z) on the device.  The challenges come from the
// transcript.  open_ok: GWC, every batch polynomial evaluates at its point to
// the v-combination of the evaluations; Shplonk, L(u) = 0 -- each holds only
// if every evaluation and every division is right.
hipError_t open_phase(Job &j, Transcript &tr, const void *hpoly, bool shplonk, bool batched, bool tamper,
                      uint8_t *out, int &ncom, int &open_ok, int &nevals, hipStream_t st) {
    using namespace zkh;
    const Pk &pk = *j.pk;
    const Shape &s = pk.sh;
    const int n = (int)s.n;
    const size_t nb = (size_t)s.n * FR;
    // n-size work slots in the extended buffers (free after the vanishing step)
    std::vector<void *> slot;
    for (void *b : {j.l_acc, j.r_acc, j.aux0, j.aux1})
        for (int q = 0; q < (1 << kExtBits); q++) slot.push_back((uint8_t *)b + (size_t)q * nb);
    void *ev_scratch = slot[15], *kate_tmp = slot[14];
    uint8_t *oc = (uint8_t *)j.open_cst;
    auto ocs = [&](int i) -> void * { return oc + (size_t)i * FR; };
    const OpenLayout lay = open_layout(j, hpoly);
    const int npoly = (int)lay.poly.size();

    // points: x w^-(blinding + 1), x, x w (the rotations last, cur, next)
    PCHECK(zk_fr_op(j.cf(J_XW), j.cf(J_X), pk.cf(C_WN), OP_MUL, 1, st));
    PCHECK(hipMemcpyAsync(j.cf(J_XWU), j.cf(J_X), FR, hipMemcpyDeviceToDevice, st));
    for (int i = 0; i <= s.blind; i++) PCHECK(zk_fr_op(j.cf(J_XWU), j.cf(J_XWU), pk.cf(C_WNINV), OP_MUL, 1, st));
    const void *pts_dev[ROT_COUNT] = {j.cf(J_XWU), j.cf(J_X), j.cf(J_XW)};
    const int rot_value[ROT_COUNT] = {-(s.blind + 1), 0, 1};

    // evaluations in the reference's order
    std::vector<int> eidx((size_t)npoly * ROT_COUNT, -1);
    const int e = (int)lay.evals.size();
    // one batched evaluation per point, each result written to its slot
    for (int r = 0; r < ROT_COUNT; r++) {
        std::vector<const void *> polys;
        std::vector<void *> outs;
        for (int i = 0; i < e; i++) {
            if (lay.evals[i].second != r) continue;
            polys.push_back(lay.poly[lay.evals[i].first]);
            outs.push_back((uint8_t *)j.evals + (size_t)i * FR);
        }
        if (polys.empty()) continue;
        if (zk_poly_eval_batch_scratch_bytes(n, (int)polys.size()) > nb) return hipErrorInvalidValue;
        PCHECK(zk_poly_eval_batch(polys.data(), outs.data(), (int)polys.size(), pts_dev[r], n, ev_scratch, st));
    }
    for (int i = 0; i < e; i++) eidx[(size_t)lay.evals[i].first * ROT_COUNT + lay.evals[i].second] = i;
    nevals = e;
    if (tamper)  // self-test of the opening check: one written evaluation off by one
        PCHECK(zk_fr_op((uint8_t *)j.evals + (size_t)(e - 1) * FR, (uint8_t *)j.evals + (size_t)(e - 1) * FR,
                        pk.cf(C_ONE), OP_ADD, 1, st));
    std::vector<Fr> ev(e);
    Fr P[ROT_COUNT];
    PCHECK(hipMemcpyAsync(ev.data(), j.evals, (size_t)e * FR, hipMemcpyDeviceToHost, st));
    for (int q = 0; q < ROT_COUNT; q++) PCHECK(hipMemcpyAsync(&P[q], pts_dev[q], FR, hipMemcpyDeviceToHost, st));
    PCHECK(hipStreamSynchronize(st));
    // every evaluation but h(x) is written (src/lib.rs:1083-1085); GWC then
    // draws v, Shplonk y' then v
    tr.scalars(ev.data() + 1, e - 1, true);
    j.last_evals = ev;
    auto eval_of = [&](int p, int r) { return ev[(size_t)eidx[(size_t)p * ROT_COUNT + r]]; };
    for (const auto &q : lay.queries)
        if (eidx[(size_t)q.first * ROT_COUNT + q.second] < 0) return hipErrorUnknown;  // a query without its evaluation
    Fr v, y = fr_zero();
    if (shplonk) {
        y = tr.challenge();
        v = tr.challenge();
        const Fr yv[2] = {v, y};  // J_V, J_SY
        PCHECK(put_challenges(j, J_V, yv, 2, st));
    } else {
        v = tr.challenge();
        PCHECK(put_challenges(j, J_V, &v, 1, st));
    }
    j.last_open_ch[0] = v;
    j.last_open_ch[1] = y;
    j.last_open_ch[2] = fr_zero();

    if (!shplonk) {
        // construct_intermediate_sets: queries grouped by rotation, groups in
        // rotation order; a group's batch is sum_i v^(len-1-i) p_i
        std::vector<const void *> W;
        std::vector<Fr> expect;
        for (int r = 0; r < ROT_COUNT; r++) {
            std::vector<int> grp;
            for (const auto &q : lay.queries)
                if (q.second == r) grp.push_back(q.first);
            if (grp.empty()) continue;
            void *B = slot[W.size()];
            Fr acc = fr_zero();
            std::vector<const void *> gp;
            for (size_t i = 0; i < grp.size(); i++) {
                gp.push_back(lay.poly[grp[i]]);
                acc = fr_add(fr_mul(acc, v), eval_of(grp[i], r));
            }
            PCHECK(zk::poly_horner_combine(B, gp.data(), (int)gp.size(), j.cf(J_V), n, st));
            PCHECK(zk_kate_divide(B, B, pts_dev[r], ocs((int)W.size()), n, kate_tmp, st));
            expect.push_back(acc);
            W.push_back(B);
        }
        const int ng = (int)W.size();
        PCHECK(commit(j, pk.g, W, out + (size_t)ncom * 64, batched, st));
        tr.points(out + (size_t)ncom * 64, ng, true);
        ncom += ng;
        std::vector<Fr> got(ng);
        PCHECK(hipMemcpyAsync(got.data(), oc, (size_t)ng * FR, hipMemcpyDeviceToHost, st));
        PCHECK(hipStreamSynchronize(st));
        open_ok = 1;
        for (int q = 0; q < ng; q++) open_ok &= fr_eq(got[q], expect[q]) ? 1 : 0;
        return hipSuccess;
    }

    // Shplonk.  Each polynomial's rotation set, polynomials in order of first
    // appearance among the queries; sets ordered as BTreeSet<Rotation>
    // (lexicographic over the sorted rotations)
    std::vector<int> mask(npoly, 0), order;
    for (const auto &q : lay.queries) {
        if (!mask[q.first]) order.push_back(q.first);
        mask[q.first] |= 1 << q.second;
    }
    auto key_of = [&](int mk) {
        std::vector<int> k;
        for (int r = 0; r < ROT_COUNT; r++)
            if (mk >> r & 1) k.push_back(rot_value[r]);
        return k;
    };
    std::vector<std::pair<std::vector<int>, int>> keys;  // (sorted rotations, mask)
    for (int p : order) {
        bool seen = false;
        for (auto &k : keys) seen |= k.second == mask[p];
        if (!seen) keys.push_back({key_of(mask[p]), mask[p]});
    }
    std::sort(keys.begin(), keys.end());
    const int nsets = (int)keys.size();
    if (nsets > 4) return hipErrorNotSupported;
    int super_mask = 0;
    for (auto &k : keys) super_mask |= k.second;
    std::vector<int> super_rots;  // the super point set, in rotation order
    for (int r = 0; r < ROT_COUNT; r++)
        if (super_mask >> r & 1) super_rots.push_back(r);
    // staging: [8, 11) super points, [12 + 3 s) diffs of set s, [24 + 3 s) r_s
    // coefficients, [36 + s) r_s(u), [40 + s) z_s, [44] Z_T(u), [45] 1/z_0
    std::vector<Fr> stage(kOpenCst, fr_zero());
    for (size_t i = 0; i < super_rots.size(); i++) stage[8 + i] = P[super_rots[i]];
    std::vector<void *> V(nsets);
    std::vector<std::vector<Fr>> r(nsets);
    std::vector<std::vector<int>> set_rots(nsets), diffs(nsets);
    for (int si = 0; si < nsets; si++) {
        for (int rr = 0; rr < ROT_COUNT; rr++) {
            if (keys[si].second >> rr & 1)
                set_rots[si].push_back(rr);
            else if (super_mask >> rr & 1)
                diffs[si].push_back(rr);
        }
        for (size_t d = 0; d < diffs[si].size(); d++) stage[12 + 3 * si + d] = P[diffs[si][d]];
        V[si] = slot[si];
        const int npts = (int)set_rots[si].size();
        std::vector<Fr> ec(npts, fr_zero()), pts(npts);
        for (int q = 0; q < npts; q++) pts[q] = P[set_rots[si][q]];
        std::vector<const void *> sp;
        for (int p : order) {
            if (mask[p] != keys[si].second) continue;
            sp.push_back(lay.poly[p]);
            for (int q = 0; q < npts; q++) ec[q] = fr_add(fr_mul(ec[q], y), eval_of(p, set_rots[si][q]));
        }
        if (!sp.empty()) PCHECK(zk::poly_horner_combine(V[si], sp.data(), (int)sp.size(), j.cf(J_SY), n, st));
        r[si] = lagrange(pts.data(), ec.data(), npts);
        for (int c = 0; c < npts; c++) stage[24 + 3 * si + c] = r[si][c];
    }
    PCHECK(hipMemcpyAsync(ocs(8), stage.data() + 8, 28 * FR, hipMemcpyHostToDevice, st));
    PCHECK(hipStreamSynchronize(st));  // the staged host values
    void *hx = slot[4], *pb = slot[5], *tmp = slot[6], *fz = slot[7];
    PCHECK(zk::fill_zero(hx, nb, st));
    for (int si = 0; si < nsets; si++) {
        PCHECK(hipMemcpyAsync(pb, V[si], nb, hipMemcpyDeviceToDevice, st));
        PCHECK(field_op(pb, pb, 0, nullptr, ocs(24 + 3 * si), 0, nullptr, (int)set_rots[si].size(), FO_SUB, st));
        bool sw = false;
        PCHECK(ntt(pb, tmp, nullptr, pk.w_n, nullptr, s.k, 8, &sw, st));
        PCHECK(zk_shplonk_h_x_merge_on(hx, j.cf(J_V), sw ? tmp : pb, pk.w_n, ocs(12 + 3 * si),
                                       (int)diffs[si].size(), n, st));
    }
    PCHECK(zk_shplonk_h_x_div_points_on(hx, pk.w_n, ocs(8), (int)super_rots.size(), n, st));
    {
        bool sw = false;
        PCHECK(ntt(hx, tmp, nullptr, pk.w_n_inv, pk.cf(C_NINV), s.k, 8, &sw, st));
        if (sw) PCHECK(hipMemcpyAsync(hx, tmp, nb, hipMemcpyDeviceToDevice, st));
    }
    PCHECK(commit(j, pk.g, std::vector<const void *>{hx}, out + (size_t)ncom * 64, batched, st));
    tr.points(out + (size_t)ncom * 64, 1, true);
    ncom += 1;
    const Fr u = tr.challenge();
    PCHECK(put_challenges(j, J_U, &u, 1, st));
    j.last_open_ch[2] = u;
    Fr zt = fr_one();
    for (int q : super_rots) zt = fr_mul(zt, fr_sub(u, P[q]));
    for (int si = 0; si < nsets; si++) {
        Fr ru = fr_zero();  // r_s(u)
        for (int c = (int)r[si].size() - 1; c >= 0; c--) ru = fr_add(fr_mul(ru, u), r[si][c]);
        Fr zs = fr_one();   // product over the points outside the set
        for (int q : diffs[si]) zs = fr_mul(zs, fr_sub(u, P[q]));
        stage[36 + si] = ru;
        stage[40 + si] = zs;
    }
    stage[44] = zt;
    stage[45] = fr_inv(stage[40]);
    PCHECK(hipMemcpyAsync(ocs(36), stage.data() + 36, 10 * FR, hipMemcpyHostToDevice, st));
    PCHECK(hipStreamSynchronize(st));  // the staged scalars
    // L(X) = (sum_s v^. z_s (V_s - r_s(u)) - Z_T(u) h(X)) / z_0, then / (X - u)
    PCHECK(zk::fill_zero(fz, nb, st));
    for (int si = 0; si < nsets; si++) {
        PCHECK(zk_fr_op(V[si], V[si], ocs(36 + si), OP_SUB, 1, st));
        PCHECK(field_op(fz, fz, 0, j.cf(J_V), V[si], 0, ocs(40 + si), n, FO_ADD, st));
    }
    PCHECK(field_op(fz, fz, 0, nullptr, hx, 0, ocs(44), n, FO_SUB, st));
    PCHECK(field_op(fz, fz, 0, ocs(45), nullptr, 0, nullptr, n, FO_LEFT, st));
    PCHECK(zk_kate_divide(fz, fz, j.cf(J_U), ocs(0), n, kate_tmp, st));
    PCHECK(commit(j, pk.g, std::vector<const void *>{fz}, out + (size_t)ncom * 64, batched, st));
    tr.points(out + (size_t)ncom * 64, 1, true);
    ncom += 1;
    Fr lu;
    PCHECK(hipMemcpyAsync(&lu, ocs(0), FR, hipMemcpyDeviceToHost, st));
    PCHECK(hipStreamSynchronize(st));
    open_ok = fr_is_zero(lu) ? 1 : 0;
    return hipSuccess;
}

}  // namespace

extern "C" {

// Device bytes a prover-core key (fixed columns, SRS stand-ins, tables) and
// one in-flight proof of this circuit hold.
size_t zkp_pk_bytes(const zkp_circuit *c) {
    Pk pk;
    if (!make_shape(c, pk.sh)) return 0;
    Arena a;
    pk_layout(pk, a);
    return a.off + 256;
}
size_t zkp_job_bytes(const zkp_circuit *c) {
    Pk pk;
    if (!make_shape(c, pk.sh)) return 0;
    Job j;
    j.pk = &pk;
    Arena a;
    job_layout(j, a);
    return a.off + 256;
}

// Builds the shape-level data once (stays resident for every proof of it).
hipError_t zkp_pk_create(const zkp_circuit *c, unsigned long long seed, hipStream_t stream, void **pk_out) {
    return zk::guard([&]() -> hipError_t {
        if (!pk_out) return hipErrorInvalidValue;
        Pk *pk = new Pk();
        if (!make_shape(c, pk->sh)) {
            delete pk;
            return hipErrorInvalidValue;
        }
        Arena dry;
        pk_layout(*pk, dry);
        pk->bytes = dry.off + 256;
        hipError_t e = hipMalloc((void **)&pk->mem, pk->bytes);
        if (e != hipSuccess) {
            delete pk;
            return e;
        }
        Arena a;
        a.base = pk->mem;
        e = zk::guard([&] {  // layout and build free below if they throw
            pk_layout(*pk, a);
            return pk_build(*pk, seed, stream);
        });
        if (e != hipSuccess) {
            (void)hipFree(pk->mem);
            delete pk;
            return e;
        }
        *pk_out = pk;
        return hipSuccess;
    });
}

hipError_t zkp_pk_srs_secret(void *pk, void *tau_out) {
    if (!pk || !tau_out) return hipErrorInvalidValue;
    return hipMemcpy(tau_out, ((Pk *)pk)->cf(C_TAU), FR, hipMemcpyDeviceToHost);
}

int zkp_pk_fixed_commitments(void *pk, void *out, int max) {
    if (!pk) return -1;
    const Pk &p = *(const Pk *)pk;
    const int count = (int)(p.vk.size() / 64);
    if (!out || max < count) return -1;
    memcpy(out, p.vk.data(), p.vk.size());
    return count;
}

int zkp_proof_openings(void *job, void *evals, int max, void *challenges3) {
    if (!job) return -1;
    const Job &j = *(const Job *)job;
    const int count = (int)j.last_evals.size();
    if (evals) {
        if (max < count) return -1;
        memcpy(evals, j.last_evals.data(), (size_t)count * FR);
    }
    if (challenges3) memcpy(challenges3, j.last_open_ch, 3 * FR);
    return count;
}

hipError_t zkp_pk_destroy(void *pk) {
    if (!pk) return hipSuccess;
    Pk *p = (Pk *)pk;
    hipError_t e = hipFree(p->mem);
    delete p;
    return e;
}

// Commitments a proof produces at most: instance (absorbed, not written),
// advice, m, permutation z, logup z, shuffle z, the random polynomial,
// quotient pieces, then the opening (GWC: one per distinct rotation;
// Shplonk: h and L).
int zkp_commitment_count(const zkp_circuit *c) {
    Shape s;
    if (!make_shape(c, s)) return 0;
    const int open = open_gwc_points(s) > 2 ? open_gwc_points(s) : 2;
    return s.I + s.A + 2 * s.L + s.Sp + s.S + 1 + kQuot + open;
}

// The last proof's bytes: every written point (compressed, 32 bytes) and
// scalar (32 bytes) in transcript order.  Returns the length; copies it when
// out has room for it (max bytes), else returns -1.  out = null: length only.
long long zkp_proof_bytes(void *job, void *out, long long max) {
    if (!job) return -1;
    const Job &j = *(const Job *)job;
    const long long len = (long long)j.proof.size();
    if (out) {
        if (max < len) return -1;
        memcpy(out, j.proof.data(), j.proof.size());
    }
    return len;
}

// The last proof's instance columns (Lagrange values, n Fr each, Montgomery):
// the public input the verifier commits itself.  rows: room per column.
hipError_t zkp_proof_instance(void *job, void *out, int rows) {
    if (!job || !out) return hipErrorInvalidValue;
    const Job &j = *(const Job *)job;
    const Shape &s = j.pk->sh;
    if (rows < (int)s.n) return hipErrorInvalidValue;
    for (int c = 0; c < s.I; c++)
        PCHECK(hipMemcpy((uint8_t *)out + (size_t)c * rows * FR, j.inst_lag[c], (size_t)s.n * FR,
                            hipMemcpyDeviceToHost));
    return hipSuccess;
}

// The verifying key's transcript representation (Montgomery Fr), the
// first scalar every proof absorbs.
hipError_t zkp_pk_vk_repr(void *pk, void *out32) {
    if (!pk || !out32) return hipErrorInvalidValue;
    memcpy(out32, &((const Pk *)pk)->vk_repr, FR);
    return hipSuccess;
}

hipError_t zkp_job_create(void *pk, void **job_out) {
    return zk::guard([&]() -> hipError_t {
        if (!pk || !job_out) return hipErrorInvalidValue;
        Job *j = new Job();
        j->pk = (const Pk *)pk;
        Arena dry;
        job_layout(*j, dry);
        j->bytes = dry.off + 256;
        hipError_t e = hipMalloc((void **)&j->mem, j->bytes);
        if (e != hipSuccess) {
            delete j;
            return e;
        }
        Arena a;
        a.base = j->mem;
        e = zk::guard([&] {
            job_layout(*j, a);
            return hipSuccess;
        });
        if (e != hipSuccess) {
            (void)hipFree(j->mem);
            delete j;
            return e;
        }
        *job_out = j;
        return hipSuccess;
    });
}

hipError_t zkp_job_destroy(void *job) {
    if (!job) return hipSuccess;
    Job *j = (Job *)job;
    if (j->h_graph) (void)hipGraphExecDestroy(j->h_graph);
    if (j->capture) (void)hipStreamDestroy(j->capture);
    if (j->upload) (void)hipStreamSynchronize(j->upload);
    for (hipEvent_t e : j->upload_ev) (void)hipEventDestroy(e);
    if (j->upload) (void)hipStreamDestroy(j->upload);
    hipError_t e = hipFree(j->mem);
    delete j;
    return e;
}

}  // extern "C"

namespace zk {
hipError_t job_set_commit_fn(void *job, CommitFn fn, void *ctx) {
    if (!job) return hipErrorInvalidValue;
    Job &j = *(Job *)job;
    j.commit_fn = fn;
    j.commit_ctx = ctx;
    return hipSuccess;
}
}  // namespace zk

extern "C" {

// Host-resident witness for the next ZKP_HOST_WITNESS proof: advice and
// instance columns (2^k Fr each, Lagrange values, Montgomery form) in host
// memory, as create_proof_from_advices receives them (src/lib.rs:205-257).
// The pointers are kept, not copied, and must stay valid until that proof
// returns; pinned memory (zk_pinned_alloc) lets the uploads overlap the
// commitments.
hipError_t zkp_set_host_witness(void *job, const void *const *advice, int advice_count, const void *const *instance,
                                int instance_count) {
    return zk::guard([&]() -> hipError_t {
        if (!job) return hipErrorInvalidValue;
        Job &j = *(Job *)job;
        const Shape &s = j.pk->sh;
        if (advice_count != s.A || instance_count != s.I || (s.A && !advice) || (s.I && !instance))
            return hipErrorInvalidValue;
        for (int c = 0; c < s.A; c++)
            if (!advice[c]) return hipErrorInvalidValue;
        for (int c = 0; c < s.I; c++)
            if (!instance[c]) return hipErrorInvalidValue;
        j.host_adv.assign(advice, advice + s.A);
        j.host_inst.assign(instance, instance + s.I);
        return hipSuccess;
    });
}

// The last proof's witness (advice then instance columns, 2^k Fr each) into
// host memory.
hipError_t zkp_proof_witness(void *job, void *const *advice, void *const *instance) {
    if (!job) return hipErrorInvalidValue;
    const Job &j = *(const Job *)job;
    const Shape &s = j.pk->sh;
    for (int c = 0; c < s.A; c++) {
        if (!advice || !advice[c]) return hipErrorInvalidValue;
        PCHECK(hipMemcpy(advice[c], j.adv_lag[c], (size_t)s.n * FR, hipMemcpyDeviceToHost));
    }
    for (int c = 0; c < s.I; c++) {
        if (!instance || !instance[c]) return hipErrorInvalidValue;
        PCHECK(hipMemcpy(instance[c], j.inst_lag[c], (size_t)s.n * FR, hipMemcpyDeviceToHost));
    }
    return hipSuccess;
}

// One proof's device arithmetic for the given seed on `stream`.  commitments
// (host, zkp_commitment_count * 64 B) receives the affine commitments in
// phase order; h_pieces (host, n Fr, may be null) the folded quotient.
hipError_t zkp_prove(void *job, unsigned long long seed, int flags, void *commitments, void *h_pieces,
                     zkp_proof_info *info, hipStream_t stream) {
    return zk::guard([&]() -> hipError_t {
        if (!job || !commitments) return hipErrorInvalidValue;
        const int batched = flags & ZKP_BATCHED;
        Job &j = *(Job *)job;
        const Pk &pk = *j.pk;
        const Shape &s = pk.sh;
        hipStream_t st = stream;
        j.seed = seed;
        j.proof.clear();
        zkp_proof_info inf;
        memset(&inf, 0, sizeof(inf));
        uint8_t *out = (uint8_t *)commitments;
        int ncom = 0;
        double t0 = now_ms(), tp = t0;
        auto mark = [&](int ph) -> hipError_t {
            PCHECK(hipStreamSynchronize(st));
            double t = now_ms();
            inf.phase_ms[ph] = t - tp;
            tp = t;
            return hipSuccess;
        };

        // Fiat-Shamir: commitments and evaluations are absorbed in the order the
        // reference writes them; every challenge is squeezed from the transcript,
        // which starts from the verifying key (vk.hash_into, src/lib.rs:242)
        Transcript tr;
        tr.proof = &j.proof;
        tr.scalars(&pk.vk_repr, 1, false);

        // ---- 0. witness (the caller's advice and instances in the reference) ----
        const bool from_host = (flags & ZKP_HOST_WITNESS) != 0;
        if (from_host && ((int)j.host_adv.size() != s.A || (int)j.host_inst.size() != s.I))
            return hipErrorInvalidValue;
        PCHECK(witness(j, from_host, st));
        if ((flags & ZKP_TAMPER_GATE) && !from_host) {
            // self-test of the quotient check: one usable row breaks the gate
            uint8_t *d = (uint8_t *)j.adv_lag[s.col_gate + 3] + (size_t)(s.u / 2) * FR;
            PCHECK(zk_fr_op(d, d, pk.cf(C_ONE), OP_ADD, 1, st));
        }
        PCHECK(mark(0));
        const double t_start = tp;

        // ---- 1. advice and instance commitments (Lagrange basis) and iNTT -------
        // one batch over advice then instances (batch_msm_and_intt_ext,
        // src/lib.rs:320-379); the instance commitments are absorbed first and
        // not written (common_point), the advice ones written (:380-385).  A
        // host-resident witness is uploaded on the job's upload stream, one
        // commitment group ahead of the MSMs that read it.
        {
            std::vector<const void *> polys(j.adv_lag.begin(), j.adv_lag.end());
            for (auto p : j.inst_lag) polys.push_back(p);
            if (from_host) PCHECK(upload_witness(j, polys.size(), st));
            std::vector<uint8_t> com(polys.size() * 64);
            PCHECK(commit(j, pk.g_lag, polys, com.data(), batched != 0, st, from_host));
            memcpy(out, com.data() + (size_t)s.A * 64, (size_t)s.I * 64);
            memcpy(out + (size_t)s.I * 64, com.data(), (size_t)s.A * 64);
            tr.points(out, s.I, false);
            tr.points(out + (size_t)s.I * 64, s.A, true);
            ncom += s.I + s.A;
            (void)tr.challenge();  // theta (src/lib.rs:395): this circuit's lookups are single columns
            PCHECK(intt_all(j, j.adv_lag, j.adv_coef, j.adv_tmp, st));
            PCHECK(intt_all(j, j.inst_lag, j.inst_coef, j.inst_tmp, st));
        }
        PCHECK(mark(1));

        // ---- 2. lookup multiplicities -------------------------------------------
        // as src/lib.rs:404-504: index the table, then count each lookup's input
        // (one table column serves every lookup of this circuit); the grand-product
        // buffers are free until phase 3 and hold the index and the counters
        if (s.L)
            PCHECK(prepare_table_lookup(pk.t_lag, nullptr, nullptr, nullptr, nullptr, j.w_temp, nullptr, 0, s.k, s.u,
                                        st));
        for (int l = 0; l < s.L; l++) {
            PCHECK(zk::wit_zero_rows(j.m_lag[l], 0, s.n, st));
            PCHECK(calc_m(j.m_lag[l], pk.t_lag, j.adv_lag[s.col_lookup + l], nullptr, j.w_temp, j.w_sum, j.w_tab,
                          nullptr, 0, s.k, s.u, st));
        }
        if (s.L) {
            std::vector<const void *> polys(j.m_lag.begin(), j.m_lag.end());
            PCHECK(commit(j, pk.g_lag, polys, out + (size_t)ncom * 64, batched != 0, st));
            tr.points(out + (size_t)ncom * 64, s.L, true);
            ncom += s.L;
            PCHECK(intt_all(j, j.m_lag, j.m_coef, j.m_tmp, st));
        }
        {
            const zkh::Fr bg[2] = {tr.challenge(), tr.challenge()};  // beta, gamma (src/lib.rs:531-532)
            PCHECK(put_challenges(j, J_BETA, bg, 2, st));
            memcpy(inf.challenges[0], &bg[0], FR);
            memcpy(inf.challenges[1], &bg[1], FR);
        }
        PCHECK(mark(2));

        // ---- 3. grand products ----------------------------------------------------
        void *beta = j.cf(J_BETA), *gamma = j.cf(J_GAMMA);
        // every lookup reads the same table: (t + beta)^-1 once per proof
        // instead of a table copy and its inversion per eval_logup_z_compose
        if (s.L > 0) PCHECK(zk::batch_invert_shifted(j.w_tab, pk.t_lag, beta, j.w_temp, (int)s.n, st));
        for (int l = 0; l < s.L; l++) {
            PCHECK(logup_sum_input_inv(j.w_sum, j.adv_lag[s.col_lookup + l], j.w_temp, beta, 0, (int)s.n, st));
            PCHECK(zk::fill_zero(j.cf(J_LASTZ), FR, st));
            PCHECK(zk::logup_z_inverted(j.zl_lag[l], j.w_sum, j.w_tab, j.m_lag[l], j.cf(J_LASTZ), (int)s.u, (int)s.n,
                                        st));
        }
        // device pointer arrays: [values | sigmas | zs], then the gate term list,
        // its rotations, the permutation sets, the permutation columns' cosets and
        // the shuffle inputs' / tables' cosets; shuffle z inputs reuse the start
        void **pa = j.ptrs;
        void **pa_sets = pa + 2 * s.P + s.Sp;
        void **pa_pv = pa_sets + s.Sp, **pa_ps = pa_pv + s.P;
        void **pa_sh = pa_ps + s.P;
        // permutation columns: instance first, then the advice ones
        auto perm_col = [&](int c, bool coef) -> void * {
            if (c < s.I) return coef ? j.inst_coef[c] : j.inst_lag[c];
            return coef ? j.adv_coef[s.col_perm + c - s.I] : j.adv_lag[s.col_perm + c - s.I];
        };
        if (s.P > 0) {
            std::vector<void *> host(2 * s.P + s.Sp);
            for (int c = 0; c < s.P; c++) {
                host[c] = perm_col(c, false);
                host[s.P + c] = pk.sig_lag[c];
            }
            for (int t = 0; t < s.Sp; t++) host[2 * s.P + t] = j.zp_lag[t];
            PCHECK(hipMemcpyAsync(pa, host.data(), host.size() * 8, hipMemcpyHostToDevice, st));
            for (int t = 0; t < s.Sp; t++) {
                int c0 = t * kChunk, nc = s.P - c0 < kChunk ? s.P - c0 : kChunk;
                PCHECK(zk_permutation_z_set(j.zp_lag[t], pa + c0, pa + s.P + c0, nc, beta, gamma, pk.cf(C_DELTAS + c0),
                                            pk.cf(C_DELTA), pk.cf(C_WN), j.w_temp, (int)s.n, st));
            }
            PCHECK(zk_permutation_chain(pa + 2 * s.P, s.Sp, (int)s.u, (int)s.n, j.w_sum, st));
            PCHECK(hipStreamSynchronize(st));  // the host pointer array
        }
        for (int t = 0; t < s.S; t++) {
            void *host[2] = {j.adv_lag[s.col_shuffle + 2 * t], j.adv_lag[s.col_shuffle + 2 * t + 1]};
            PCHECK(hipMemcpyAsync(pa, host, 16, hipMemcpyHostToDevice, st));
            PCHECK(zk_shuffle_z(j.zs_lag[t], pa, pa + 1, 1, beta, j.w_temp, (int)s.n, st));
            PCHECK(hipStreamSynchronize(st));
        }
        {
            // permutation, lookup, shuffle, then the vanishing argument's random
            // polynomial: the reference's transcript order (src/lib.rs:968-980)
            std::vector<const void *> polys;
            for (auto p : j.zp_lag) polys.push_back(p);
            for (auto p : j.zl_lag) polys.push_back(p);
            for (auto p : j.zs_lag) polys.push_back(p);
            polys.push_back(j.rnd_lag);
            PCHECK(commit(j, pk.g_lag, polys, out + (size_t)ncom * 64, batched != 0, st));
            tr.points(out + (size_t)ncom * 64, (int)polys.size(), true);
            ncom += (int)polys.size();
            const zkh::Fr yc = tr.challenge();  // y (src/lib.rs:988)
            PCHECK(put_challenges(j, J_Y, &yc, 1, st));
            memcpy(inf.challenges[2], &yc, FR);
            // closure of every grand product at row u: logup 0, permutation (last
            // set) 1, shuffle 1
            int bad = 0;  // bit 0: a logup sum, bit 1: the permutation, bit 2: a shuffle
            uint64_t one[4];
            std::vector<uint64_t> got((size_t)(s.L + s.S + 1) * 4, 0);
            PCHECK(hipMemcpyAsync(one, pk.cf(C_ONE), FR, hipMemcpyDeviceToHost, st));
            int gi = 0;
            for (auto z : j.zl_lag)
                PCHECK(hipMemcpyAsync(&got[(size_t)4 * gi++], (const uint8_t *)z + (size_t)s.u * FR, FR,
                                      hipMemcpyDeviceToHost, st));
            for (auto z : j.zs_lag)
                PCHECK(hipMemcpyAsync(&got[(size_t)4 * gi++], (const uint8_t *)z + (size_t)s.u * FR, FR,
                                      hipMemcpyDeviceToHost, st));
            if (s.Sp)
                PCHECK(hipMemcpyAsync(&got[(size_t)4 * gi], (const uint8_t *)j.zp_lag[s.Sp - 1] + (size_t)s.u * FR,
                                      FR, hipMemcpyDeviceToHost, st));
            PCHECK(hipStreamSynchronize(st));
            for (int l = 0; l < s.L; l++) {
                const uint64_t *g = &got[(size_t)4 * l];
                if (g[0] | g[1] | g[2] | g[3]) bad |= 1;
            }
            for (int t = 0; t < s.S; t++)
                if (memcmp(&got[(size_t)4 * (s.L + t)], one, FR) != 0) bad |= 4;
            if (s.Sp && memcmp(&got[(size_t)4 * (s.L + s.S)], one, FR) != 0) bad |= 2;
            inf.z_ok = bad ? 0 : 1;
            inf.z_fail = bad;
            PCHECK(intt_all(j, j.zl_lag, j.zl_coef, j.zl_tmp, st));
            PCHECK(intt_all(j, j.zp_lag, j.zp_coef, j.zp_tmp, st));
            PCHECK(intt_all(j, j.zs_lag, j.zs_coef, j.zs_tmp, st));
            PCHECK(intt_all(j, {j.rnd_lag}, {j.rnd_coef}, {j.rnd_tmp}, st));
        }
        PCHECK(mark(3));

        // ---- 4. h(X) on the extended coset ----------------------------------------
        // ext slots: gate a,b,c,d | f_l | m_l | zl_l | perm columns (instance,
        // advice) | zp | shuffle in, sh | zs | coefficient-gate outputs e_j.
        // Every buffer of the round is fixed for the job, so its device
        // pointer arrays are staged once, and with ZKP_GRAPH the whole round
        // (coset extensions and constraint kernels, nothing but launches) is
        // captured on the job's first proof and replayed on later ones.
        const int E_GATE = 0, E_F = 4, E_M = E_F + s.L, E_ZL = E_M + s.L, E_V = E_ZL + s.L, E_ZP = E_V + s.P,
                  E_IN = E_ZP + s.Sp, E_ZS = E_IN + 2 * s.S, E_EG = E_ZS + s.S;
        void *y = j.cf(J_Y);
        const int rot1 = 1 << kExtBits;  // w X on the coset
        const int N = (int)s.N;
        // the gates, folded in y: y^FG q (a b + c - d) + sum_j y^(FG-1-j) q (f_j a - e_j)
        std::vector<const void *> gate_f(s.FG), gate_e(s.FG);
        for (int g = 0; g < s.FG; g++) {
            gate_f[g] = pk.fix_ext[g];
            gate_e[g] = j.ext[E_EG + g];
        }
        if (!j.p4_staged) {
            std::vector<void *> sets(s.Sp + 2 * s.P + 2 * s.S);
            for (int t = 0; t < s.Sp; t++) sets[t] = j.ext[E_ZP + t];
            for (int c = 0; c < s.P; c++) {
                sets[s.Sp + c] = j.ext[E_V + c];
                sets[s.Sp + s.P + c] = pk.sig_ext[c];
            }
            for (int t = 0; t < s.S; t++) {
                sets[s.Sp + 2 * s.P + t] = j.ext[E_IN + 2 * t];
                sets[s.Sp + 2 * s.P + s.S + t] = j.ext[E_IN + 2 * t + 1];
            }
            // pa_sets, pa_pv, pa_ps and pa_sh are consecutive
            if (!sets.empty())
                PCHECK(hipMemcpyAsync(pa_sets, sets.data(), sets.size() * 8, hipMemcpyHostToDevice, st));
            PCHECK(hipStreamSynchronize(st));  // the host arrays are temporaries
            j.p4_staged = true;
        }
        auto h_round = [&](hipStream_t st) -> hipError_t {
            std::vector<const void *> c;
            for (int i = 0; i < 4; i++) c.push_back(j.adv_coef[s.col_gate + i]);
            for (int l = 0; l < s.L; l++) c.push_back(j.adv_coef[s.col_lookup + l]);
            for (auto p : j.m_coef) c.push_back(p);
            for (auto p : j.zl_coef) c.push_back(p);
            for (int i = 0; i < s.P; i++) c.push_back(perm_col(i, true));
            for (auto p : j.zp_coef) c.push_back(p);
            for (int t = 0; t < 2 * s.S; t++) c.push_back(j.adv_coef[s.col_shuffle + t]);
            for (auto p : j.zs_coef) c.push_back(p);
            for (int g = 0; g < s.FG; g++) c.push_back(j.adv_coef[s.col_fixed_gate + g]);
            PCHECK(extend_all(j, c, 0, st));
            // h starts as the gate sum (written, not accumulated)
            PCHECK(zk::gate_eval_h(j.h, pk.q_ext, j.ext[E_GATE], j.ext[E_GATE + 1], j.ext[E_GATE + 2],
                                   j.ext[E_GATE + 3], gate_f.data(), gate_e.data(), s.FG, y, N, st));
            // permutation: p1, p2, then per set the fused product term (sigma
            // and X resident as cosets)
            if (s.Sp > 0) {
                PCHECK(zk_permutation_eval_h_p1_on(j.h, j.ext[E_ZP], j.ext[E_ZP + s.Sp - 1], pk.l0_ext, pk.llast_ext,
                                                   y, N, st));
                PCHECK(zk_permutation_eval_h_p2_on(j.h, pa_sets, pk.l0_ext, pk.llast_ext, y, s.Sp, (int)s.u * rot1, N,
                                                   st));
                for (int c0 = 0; c0 < s.P; c0++)
                    PCHECK(zk_fr_op(j.cf(J_BDELTA + c0), beta, pk.cf(C_DELTAS + c0), OP_MUL, 1, st));
                for (int t = 0; t < s.Sp; t++) {
                    int c0 = t * kChunk, nc = s.P - c0 < kChunk ? s.P - c0 : kChunk;
                    PCHECK(zk_permutation_eval_h_set(j.h, j.ext[E_ZP + t], pa_pv + c0, pa_ps + c0, nc, beta, gamma,
                                                     j.cf(J_BDELTA + c0), pk.x_ext, pk.lact_ext, y, rot1, N, st));
                }
            }
            // logup: (t + beta) once (every lookup reads the one table), (f + beta),
            // product-sum 1
            if (s.L > 0) PCHECK(field_op(j.aux1, pk.t_ext, 0, nullptr, nullptr, 0, beta, N, 0, st));
            for (int l = 0; l < s.L; l++) {
                PCHECK(field_op(j.aux0, j.ext[E_F + l], 0, nullptr, nullptr, 0, beta, N, 0, st));
                PCHECK(logup_eval_h(j.h, j.aux0, pk.ones_ext, j.aux1, j.ext[E_M + l], j.ext[E_ZL + l],
                                    j.ext[E_ZL + l], pk.l0_ext, pk.llast_ext, pk.lact_ext, y, rot1, N, st));
            }
            // shuffle: each group one plain input column and one plain table
            // column, the case the reference hands to shuffle_eval_h_v2 with
            // betas [beta, ...] (src/eval_h.rs:1095-1152); the kernel adds beta
            for (int t = 0; t < s.S; t++)
                PCHECK(zk_shuffle_eval_h_v2_on(j.h, pa_sh + t, pa_sh + s.S + t, beta, 1, j.ext[E_ZS + t], pk.l0_ext,
                                               pk.llast_ext, pk.lact_ext, y, rot1, N, st));
            return hipSuccess;
        };
        if (flags & ZKP_GRAPH) {
            if (!j.h_graph) {
                // recorded on a stream of the job's own (the caller's may be the
                // legacy stream, which cannot be captured), launched on st
                if (!j.capture) PCHECK(hipStreamCreateWithFlags(&j.capture, hipStreamNonBlocking));
                hipGraph_t g = nullptr;
                PCHECK(hipStreamBeginCapture(j.capture, hipStreamCaptureModeThreadLocal));
                hipError_t ce = h_round(j.capture);
                hipError_t ee = hipStreamEndCapture(j.capture, &g);
                if (ce == hipSuccess) ce = ee;
                if (ce == hipSuccess) ce = hipGraphInstantiate(&j.h_graph, g, nullptr, nullptr, 0);
                if (g) (void)hipGraphDestroy(g);
                if (ce != hipSuccess) {
                    j.h_graph = nullptr;
                    return ce;
                }
            }
            PCHECK(hipGraphLaunch(j.h_graph, st));
        } else {
            PCHECK(h_round(st));
        }
        PCHECK(mark(4));

        // ---- 5. vanishing -----------------------------------------------------------
        const void *hpoly = nullptr;  // h(X) folded in x^n: opened at x in phase 6
        {
            PCHECK(zk_field_mul_zip_on(j.h, pk.cf(C_TEV), 4, N, st));
            bool sw = false;
            PCHECK(ntt(j.h, j.h_tmp, nullptr, pk.w_e_inv, pk.cf(C_NNINV), s.ek, 8, &sw, st));
            void *c = sw ? j.h_tmp : j.h;
            PCHECK(extended_prepare(c, pk.cf(C_INV0), 3, (int)s.n, N, 1, st));
            // all constraints hold iff h is a polynomial of degree < 3n
            PCHECK(zk::fill_zero(j.flag, 16, st));
            PCHECK(zk::wit_any_nonzero(c, (uint64_t)kQuot * s.n, s.N, j.flag, st));
            uint32_t flag = 1;
            PCHECK(hipMemcpyAsync(&flag, j.flag, 4, hipMemcpyDeviceToHost, st));
            PCHECK(hipStreamSynchronize(st));
            inf.quotient_ok = flag == 0 ? 1 : 0;
            std::vector<const void *> pieces;
            for (int i = 0; i < kQuot; i++) pieces.push_back((uint8_t *)c + (size_t)i * s.n * FR);
            PCHECK(commit(j, pk.g, pieces, out + (size_t)ncom * 64, batched != 0, st));
            tr.points(out + (size_t)ncom * 64, kQuot, true);
            ncom += kQuot;
            const zkh::Fr xc = tr.challenge();  // x
            PCHECK(put_challenges(j, J_X, &xc, 1, st));
            memcpy(inf.challenges[3], &xc, FR);
            // x^n, then h_pieces = sum_i piece_i x^(n i) by Horner into the last piece
            PCHECK(square_times(j.cf(J_XN), j.cf(J_X), s.k, st));
            uint8_t *last = (uint8_t *)c + (size_t)(kQuot - 1) * s.n * FR;
            for (int i = kQuot - 2; i >= 0; i--)
                PCHECK(field_op(last, last, 0, j.cf(J_XN), (uint8_t *)c + (size_t)i * s.n * FR, 0, nullptr, (int)s.n, 0,
                                st));
            if (h_pieces) PCHECK(hipMemcpyAsync(h_pieces, last, (size_t)s.n * FR, hipMemcpyDeviceToHost, st));
            hpoly = last;
        }
        PCHECK(mark(5));

        // ---- 6. multi-open ------------------------------------------------------------
        if (!(flags & ZKP_NO_OPEN)) {
            int ok = 0, ne = 0;
            PCHECK(open_phase(j, tr, hpoly, (flags & ZKP_SHPLONK) != 0, batched != 0, (flags & ZKP_TAMPER_EVAL) != 0,
                              out, ncom, ok, ne, st));
            inf.open_ok = ok;
            inf.nevals = ne;
        }
        PCHECK(mark(6));
        {
            uint8_t d[64];
            zk_transcript_squeeze(tr.t, nullptr, d);
            memcpy(inf.transcript, d, sizeof(inf.transcript));
        }
        inf.total_ms = tp - t_start;
        inf.ncommit = ncom;
        if (info) *info = inf;
        return hipSuccess;
    });
}

}  // extern "C"
