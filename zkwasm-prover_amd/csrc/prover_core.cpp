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

#include <chrono>
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

struct Shape {
    int k, ek, A, F, I, L, P, S, Sp, tlog, slog, blind;
    uint32_t n, N, u, tsize, R;
    // advice column roles
    int col_gate, col_lookup, col_perm, col_shuffle;
};

bool make_shape(const zkp_circuit *c, Shape &s) {
    if (!c || c->k < 4 || c->k > 24 || c->lookups < 0 || c->perm_cols < 0 || c->shuffles < 0 || c->blinding < 1 ||
        c->table_log < 1 || c->shuffle_log < 1)
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
    s.S = c->shuffles;
    s.Sp = (s.P + kChunk - 1) / kChunk;
    s.tlog = c->table_log;
    s.slog = c->shuffle_log;
    s.tsize = (s.tlog < 31 && (1u << s.tlog) < s.u) ? (1u << s.tlog) : s.u;
    s.R = (s.slog < 31 && (1u << s.slog) < s.u) ? (1u << s.slog) : s.u;
    s.col_gate = 0;
    s.col_lookup = 4;
    s.col_perm = 4 + s.L;
    s.col_shuffle = 4 + s.L + s.P;
    s.A = c->advice;
    s.F = c->fixed < 2 ? 2 : c->fixed;
    s.I = c->instance < 0 ? 0 : c->instance;
    if (s.A < s.col_shuffle + 2 * s.S) return false;
    if (4 * s.P + 2 * s.Sp + 24 + 2 * s.S > 512) return false;  // device pointer area of a job
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
    uint8_t *cst;  // Fr constants
    void *t_lag, *t_ext, *q_ext, *l0_ext, *llast_ext, *lact_ext, *ones_ext, *x_ext;
    void *t_coef, *q_coef;  // fixed columns opened in the multi-open step
    std::vector<uint8_t> vk;  // commitments of q, t, sigmas (affine, host)
    std::vector<void *> sig_lag, sig_coef, sig_ext;
    uint32_t *perm_head, *perm_target;
    std::vector<uint32_t *> shuf_perm;
    void **sig_lag_dev;  // device array of P pointers
    void *scratch_n[3];  // setup work buffers (n)
    void *scratch_e;     // setup work buffer (N)
    void *cf(int c) const { return cst + (size_t)c * FR; }
};

void pk_layout(Pk &pk, Arena &a) {
    const Shape &s = pk.sh;
    pk.g_lag = a.take((size_t)s.n * 64);
    pk.g = a.take((size_t)s.n * 64);
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
    // the verifying key: commitments of the fixed columns the proof opens
    {
        std::vector<const void *> fixed = {pk.q_coef, pk.t_coef};
        for (auto c : pk.sig_coef) fixed.push_back(c);
        PCHECK(hipStreamSynchronize(st));
        pk.vk.assign(fixed.size() * 64, 0);
        PCHECK(zkp_batch_msm(pk.vk.data(), pk.g, fixed.data(), (int)fixed.size(), (int)s.n, 0));
    }
    return hipStreamSynchronize(st);
}

// ---- multi-open query layout -------------------------------------------------
// Polynomials grouped by the points they are opened at (points: x, x w,
// x w^u): set 0 at {x}: advice, q, t, sigmas, m, h; set 1 at {x, x w}:
// logup z, shuffle z, the last permutation z; set 2 at {x, x w, x w^u}: the
// other permutation z (u = the last usable row, the halo2 "last" rotation).
int open_set_size(const Shape &s, int set) {
    if (set == 0) return s.A + 2 + s.P + s.L + 1;
    if (set == 1) return s.L + s.S + (s.Sp > 0 ? 1 : 0);
    return s.Sp > 1 ? s.Sp - 1 : 0;
}
int open_eval_count(const Shape &s) { return open_set_size(s, 0) + 2 * open_set_size(s, 1) + 3 * open_set_size(s, 2); }
// GWC commits one quotient per distinct point, Shplonk two polynomials
int open_gwc_points(const Shape &s) {
    return open_set_size(s, 2) ? 3 : open_set_size(s, 1) ? 2 : 1;
}
constexpr int kOpenCst = 64;

// ---- per-proof state ---------------------------------------------------------
// beta, gamma, y, x; x^n; logup carry; opening challenges v, y', u; opening
// points x w, x w^u; then beta delta^j per permutation column
enum JConst {
    J_BETA, J_GAMMA, J_Y, J_X, J_XN, J_LASTZ, J_V, J_SY, J_U, J_XW, J_XWU, J_BDELTA,
    J_COUNT_FIXED = J_BDELTA
};

struct Job {
    const Pk *pk;
    uint64_t seed;
    uint8_t *mem = nullptr;
    size_t bytes = 0;
    std::vector<void *> adv_lag, adv_coef, adv_tmp;
    std::vector<void *> m_lag, m_coef, m_tmp, zl_lag, zl_coef, zl_tmp, zp_lag, zp_coef, zp_tmp, zs_lag, zs_coef,
        zs_tmp;
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
    void *cf(int c) const { return cst + (size_t)c * FR; }
};

// how many polynomials get an extended coset in phase 4
int ext_count(const Shape &s) { return 4 + 3 * s.L + s.P + s.Sp + 3 * s.S; }

size_t msm_group_bytes(const Shape &s) {
    std::vector<zk_msm_item> it(kMsmGroup, zk_msm_item{nullptr, nullptr, (int)s.n, 0});
    return zk_msm_batch_scratch_bytes(it.data(), kMsmGroup, 0);
}

void job_layout(Job &j, Arena &a) {
    const Shape &s = j.pk->sh;
    auto cols = [&](std::vector<void *> &v, int c, uint32_t len) {
        v.assign(c, nullptr);
        for (auto &p : v) p = a.fr(len);
    };
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
    j.cst = a.take((size_t)(J_COUNT_FIXED + 3 * s.P + 8) * FR);
    j.ptrs = (void **)a.take(4096);
    j.msm_bytes = msm_group_bytes(s);
    j.msm_scratch = a.take(j.msm_bytes);
    j.msm_out = a.take((size_t)kMsmGroup * ZK_MSM_OUT_BYTES);
    j.flag = (uint32_t *)a.take(256);
    j.evals = a.fr((size_t)open_eval_count(s));
    j.open_cst = a.fr(kOpenCst);
}

// commitments of polys (device Lagrange or coefficient values of n) against
// bases; out: affine points (host, 64 B each)
hipError_t commit(Job &j, const void *bases, const std::vector<const void *> &polys, uint8_t *out, bool batched,
                  hipStream_t st) {
    const Shape &s = j.pk->sh;
    for (size_t first = 0; first < polys.size(); first += kMsmGroup) {
        int cnt = (int)(polys.size() - first < (size_t)kMsmGroup ? polys.size() - first : kMsmGroup);
        if (batched) {
            std::vector<zk_msm_item> it(cnt);
            for (int m = 0; m < cnt; m++) it[m] = zk_msm_item{bases, polys[first + m], (int)s.n, 0};
            PCHECK(zk_msm_batch(j.msm_out, it.data(), cnt, 0, j.msm_scratch, j.msm_bytes, st));
        } else {
            for (int m = 0; m < cnt; m++)
                PCHECK(zk_msm((uint8_t *)j.msm_out + (size_t)m * ZK_MSM_OUT_BYTES, bases, polys[first + m],
                              (int)s.n, 0, j.msm_scratch, j.msm_bytes, st));
        }
        for (int m = 0; m < cnt; m++) {
            uint8_t jac[96];
            PCHECK(hipMemcpyAsync(jac, (uint8_t *)j.msm_out + (size_t)m * ZK_MSM_OUT_BYTES, 96,
                                  hipMemcpyDeviceToHost, st));
            PCHECK(hipStreamSynchronize(st));
            bool id = true;
            for (int b = 64; b < 96; b++) id &= jac[b] == 0;
            uint8_t *o = out + (first + m) * 64;
            if (id)
                memset(o, 0, 64);
            else
                memcpy(o, jac, 64);
        }
    }
    return hipSuccess;
}

// iNTT of each lag column into coef (tmp spare); coef[i] is updated to
// whichever buffer holds the result, tmp[i] to the other one
hipError_t intt_all(Job &j, std::vector<void *> &lag, std::vector<void *> &coef, std::vector<void *> &tmp,
                    hipStream_t st) {
    const Shape &s = j.pk->sh;
    const Pk &pk = *j.pk;
    size_t c = lag.size();
    if (!c) return hipSuccess;
    // the first pass reads the Lagrange values in place: no copy into coef
    std::vector<unsigned char> sw(c);
    PCHECK(zk::ntt_batch_from(lag.data(), coef.data(), tmp.data(), pk.w_n_inv, pk.cf(C_NINV), s.k, (int)c, sw.data(),
                              st));
    for (size_t i = 0; i < c; i++)
        if (sw[i]) std::swap(coef[i], tmp[i]);
    return hipSuccess;
}

// extended cosets of coefficient polys into j.ext[first..]; ext pointers are
// updated to where each result landed (the spare takes the other buffer)
hipError_t extend_all(Job &j, const std::vector<const void *> &coefs, int first, hipStream_t st) {
    const Shape &s = j.pk->sh;
    const Pk &pk = *j.pk;
    for (size_t g = 0; g < coefs.size(); g += kExtGroup) {
        int cnt = (int)(coefs.size() - g < (size_t)kExtGroup ? coefs.size() - g : kExtGroup);
        std::vector<zk_coset_item> it(cnt);
        for (int m = 0; m < cnt; m++)
            it[m] = zk_coset_item{coefs[g + m], j.ext[first + g + m], j.ext_tmp[m], pk.w_e, s.k, kExtBits};
        std::vector<unsigned char> sw(cnt);
        PCHECK(zk_coset_extend_batch(it.data(), cnt, pk.cf(C_FWD0), 3, sw.data(), st));
        for (int m = 0; m < cnt; m++)
            if (sw[m]) std::swap(j.ext[first + g + m], j.ext_tmp[m]);
    }
    return hipSuccess;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

hipError_t witness(Job &j, hipStream_t st) {
    const Shape &s = j.pk->sh;
    const Pk &pk = *j.pk;
    const uint64_t seed = j.seed;
    for (int c = 0; c < s.A; c++) PCHECK(zk_rand_fr(j.adv_lag[c], (int)s.n, seed * 1000003ull + c, st));
    // gate: d = a * b + c on the usable rows
    PCHECK(zk::wit_gate_rows(j.adv_lag[3], j.adv_lag[0], j.adv_lag[1], j.adv_lag[2], s.u, st));
    // lookups: inputs are table values on the usable rows
    for (int l = 0; l < s.L; l++) {
        PCHECK(zk::wit_lookup_values(j.vals, s.u, s.tsize, seed * 7919ull + l, st));
        PCHECK(zk::wit_u32_rows_to_fr(j.adv_lag[s.col_lookup + l], j.vals, s.u, st));
    }
    // copy cycles: every cell of a cycle takes its head's value
    if (s.P > 0) {
        std::vector<void *> cols(s.P);
        for (int c = 0; c < s.P; c++) cols[c] = j.adv_lag[s.col_perm + c];
        PCHECK(hipMemcpyAsync(j.ptrs, cols.data(), (size_t)s.P * 8, hipMemcpyHostToDevice, st));
        PCHECK(zk::wit_cycle_values(j.ptrs, pk.perm_head, (uint32_t)s.P, s.u, st));
        PCHECK(hipStreamSynchronize(st));  // cols is copied from host memory
    }
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
struct Transcript {
    void *t = zk_transcript_new();
    ~Transcript() { zk_transcript_free(t); }
    // affine commitments as commit() writes them (Montgomery x, y; zeros for
    // the identity, which halo2 cannot absorb: absorbed here as x = y = 0)
    void points(const uint8_t *aff, int count) {
        for (int i = 0; i < count; i++) {
            zkh::Fp x, y;
            memcpy(&x, aff + (size_t)i * 64, 32);
            memcpy(&y, aff + (size_t)i * 64 + 32, 32);
            const zkh::Fp one_raw{{1, 0, 0, 0}};
            if (!zkh::fp_is_zero(x) || !zkh::fp_is_zero(y)) {
                x = zkh::fp_mul(x, one_raw);  // canonical
                y = zkh::fp_mul(y, one_raw);
            }
            zk_transcript_common_point(t, &x, &y);
        }
    }
    void scalars(const zkh::Fr *v, int count) {
        const zkh::Fr one_raw{{1, 0, 0, 0}};
        for (int i = 0; i < count; i++) {
            const zkh::Fr c = zkh::fr_mul(v[i], one_raw);
            zk_transcript_common_scalar(t, &c);
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

// Multi-open of every query (src/multiopen.rs): evaluations of each
// polynomial at its points (batch_poly_eval, src/eval_poly.rs:18), then GWC
// (one quotient per point, :68-177) or Shplonk (h(X) and L(X), :283-561) with
// the divisions by (X - z) on the device.  The challenges come from the
// transcript.  open_ok: GWC, every batch polynomial evaluates at its point to the
// v-combination of the evaluations; Shplonk, L(u) = 0 -- each holds only if
// every evaluation and every division is right.
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
    std::vector<const void *> sets[3];
    for (auto p : j.adv_coef) sets[0].push_back(p);
    sets[0].push_back(pk.q_coef);
    sets[0].push_back(pk.t_coef);
    for (auto p : pk.sig_coef) sets[0].push_back(p);
    for (auto p : j.m_coef) sets[0].push_back(p);
    sets[0].push_back(hpoly);
    for (auto p : j.zl_coef) sets[1].push_back(p);
    for (auto p : j.zs_coef) sets[1].push_back(p);
    if (s.Sp > 0) sets[1].push_back(j.zp_coef[s.Sp - 1]);
    for (int t = 0; t + 1 < s.Sp; t++) sets[2].push_back(j.zp_coef[t]);
    const int nsets = sets[2].empty() ? (sets[1].empty() ? 1 : 2) : 3;  // = number of distinct points

    // points x, x w, x w^u with w^u = (w^-1)^(blinding + 1)
    PCHECK(zk_fr_op(j.cf(J_XW), j.cf(J_X), pk.cf(C_WN), OP_MUL, 1, st));
    PCHECK(hipMemcpyAsync(j.cf(J_XWU), j.cf(J_X), FR, hipMemcpyDeviceToDevice, st));
    for (int i = 0; i <= s.blind; i++) PCHECK(zk_fr_op(j.cf(J_XWU), j.cf(J_XWU), pk.cf(C_WNINV), OP_MUL, 1, st));
    const void *pts_dev[3] = {j.cf(J_X), j.cf(J_XW), j.cf(J_XWU)};

    // evaluations: set by set, polynomial by polynomial, point by point
    int e = 0;
    for (int set = 0; set < 3; set++)
        for (auto p : sets[set])
            for (int q = 0; q <= set; q++) {
                PCHECK(poly_eval(p, ev_scratch, pts_dev[q], n, st));
                PCHECK(hipMemcpyAsync((uint8_t *)j.evals + (size_t)e * FR, ev_scratch, FR, hipMemcpyDeviceToDevice,
                                      st));
                e++;
            }
    nevals = e;
    if (tamper)  // self-test of the opening check: one evaluation off by one
        PCHECK(zk_fr_op((uint8_t *)j.evals + (size_t)(e - 1) * FR, (uint8_t *)j.evals + (size_t)(e - 1) * FR,
                        pk.cf(C_ONE), OP_ADD, 1, st));
    std::vector<Fr> ev(e);
    Fr P[3];
    PCHECK(hipMemcpyAsync(ev.data(), j.evals, (size_t)e * FR, hipMemcpyDeviceToHost, st));
    for (int q = 0; q < 3; q++) PCHECK(hipMemcpyAsync(&P[q], pts_dev[q], FR, hipMemcpyDeviceToHost, st));
    PCHECK(hipStreamSynchronize(st));
    // the evaluations enter the transcript; GWC draws v, Shplonk y' then v
    // (src/multiopen.rs:85, :302-305); Shplonk's u follows h's commitment
    tr.scalars(ev.data(), e);
    j.last_evals = ev;
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
    // Horner in c over the evaluations of `set` at point q, in query order
    auto combine = [&](int set, int q, const Fr &c) {
        int first = 0;
        for (int t = 0; t < set; t++) first += (int)sets[t].size() * (t + 1);
        Fr acc = fr_zero();
        for (size_t i = 0; i < sets[set].size(); i++) acc = fr_add(fr_mul(acc, c), ev[first + i * (set + 1) + q]);
        return acc;
    };

    if (!shplonk) {
        // one batch per point: every query at point q, sets q.. in order
        std::vector<const void *> W;
        std::vector<Fr> expect(nsets);
        for (int q = 0; q < nsets; q++) {
            void *B = slot[q];
            bool first = true;
            Fr acc = fr_zero();
            for (int set = q; set < 3; set++) {
                Fr cs = combine(set, q, v);
                // the set's block of the Horner sum: acc v^len + (set's own sum)
                for (size_t i = 0; i < sets[set].size(); i++) acc = fr_mul(acc, v);
                acc = fr_add(acc, cs);
                for (auto p : sets[set]) {
                    if (first)
                        PCHECK(hipMemcpyAsync(B, p, nb, hipMemcpyDeviceToDevice, st));
                    else
                        PCHECK(field_op(B, B, 0, j.cf(J_V), p, 0, nullptr, n, FO_ADD, st));
                    first = false;
                }
            }
            expect[q] = acc;
            PCHECK(zk_kate_divide(B, B, pts_dev[q], ocs(q), n, kate_tmp, st));
            W.push_back(B);
        }
        PCHECK(commit(j, pk.g, W, out + (size_t)ncom * 64, batched, st));
        tr.points(out + (size_t)ncom * 64, nsets);
        ncom += nsets;
        std::vector<Fr> got(nsets);
        PCHECK(hipMemcpyAsync(got.data(), oc, (size_t)nsets * FR, hipMemcpyDeviceToHost, st));
        PCHECK(hipStreamSynchronize(st));
        open_ok = 1;
        for (int q = 0; q < nsets; q++) open_ok &= fr_eq(got[q], expect[q]) ? 1 : 0;
        return hipSuccess;
    }

    // Shplonk.  V_s = y-combination of set s; r_s interpolates its evaluations
    void *V[3] = {slot[0], slot[1], slot[2]};
    void *hx = slot[3], *pb = slot[4], *tmp = slot[5], *fz = slot[6];
    std::vector<Fr> r[3];
    for (int set = 0; set < nsets; set++) {
        bool first = true;
        for (auto p : sets[set]) {
            if (first)
                PCHECK(hipMemcpyAsync(V[set], p, nb, hipMemcpyDeviceToDevice, st));
            else
                PCHECK(field_op(V[set], V[set], 0, j.cf(J_SY), p, 0, nullptr, n, FO_ADD, st));
            first = false;
        }
        Fr ec[3];
        for (int q = 0; q <= set; q++) ec[q] = combine(set, q, y);
        r[set] = lagrange(P, ec, set + 1);
    }
    // host staging of the scalars the device reads: [0, 9) r coefficients of
    // the three sets, then (once u is drawn) per set u-evaluation and z_s,
    // then Z_T(u), 1/z_0
    std::vector<Fr> stage(kOpenCst - 32, fr_zero());
    for (int set = 0; set < nsets; set++)
        for (int c = 0; c <= set; c++) stage[3 * set + c] = r[set][c];
    PCHECK(hipMemcpyAsync(ocs(32), stage.data(), 9 * FR, hipMemcpyHostToDevice, st));
    // point arrays on the device: [8, 11) the points in order
    for (int q = 0; q < nsets; q++) PCHECK(hipMemcpyAsync(ocs(8 + q), pts_dev[q], FR, hipMemcpyDeviceToDevice, st));
    PCHECK(hipStreamSynchronize(st));  // the staged coefficients

    PCHECK(hipMemsetAsync(hx, 0, nb, st));
    for (int set = 0; set < nsets; set++) {
        PCHECK(hipMemcpyAsync(pb, V[set], nb, hipMemcpyDeviceToDevice, st));
        PCHECK(field_op(pb, pb, 0, nullptr, ocs(32 + 3 * set), 0, nullptr, set + 1, FO_SUB, st));
        bool sw = false;
        PCHECK(ntt(pb, tmp, nullptr, pk.w_n, nullptr, s.k, 8, &sw, st));
        // the points outside the set are ocs(8 + set + 1 ..)
        PCHECK(zk_shplonk_h_x_merge_on(hx, j.cf(J_V), sw ? tmp : pb, pk.w_n, ocs(8 + set + 1), nsets - set - 1, n,
                                       st));
    }
    PCHECK(zk_shplonk_h_x_div_points_on(hx, pk.w_n, ocs(8), nsets, n, st));
    {
        bool sw = false;
        PCHECK(ntt(hx, tmp, nullptr, pk.w_n_inv, pk.cf(C_NINV), s.k, 8, &sw, st));
        if (sw) PCHECK(hipMemcpyAsync(hx, tmp, nb, hipMemcpyDeviceToDevice, st));
    }
    PCHECK(commit(j, pk.g, std::vector<const void *>{hx}, out + (size_t)ncom * 64, batched, st));
    tr.points(out + (size_t)ncom * 64, 1);
    ncom += 1;
    const Fr u = tr.challenge();
    PCHECK(put_challenges(j, J_U, &u, 1, st));
    j.last_open_ch[2] = u;
    Fr zt = fr_one(), z0inv = fr_one();
    for (int q = 0; q < nsets; q++) zt = fr_mul(zt, fr_sub(u, P[q]));
    for (int set = 0; set < nsets; set++) {
        Fr ru = fr_zero();  // r_s(u)
        for (int c = set; c >= 0; c--) ru = fr_add(fr_mul(ru, u), r[set][c]);
        Fr zs = fr_one();   // prod over the points outside the set
        for (int q = set + 1; q < nsets; q++) zs = fr_mul(zs, fr_sub(u, P[q]));
        stage[9 + 2 * set] = ru;
        stage[10 + 2 * set] = zs;
        if (set == 0) z0inv = fr_inv(zs);
    }
    stage[15] = zt;
    stage[16] = z0inv;
    PCHECK(hipMemcpyAsync(ocs(32 + 9), stage.data() + 9, (stage.size() - 9) * FR, hipMemcpyHostToDevice, st));
    PCHECK(hipStreamSynchronize(st));  // the staged scalars
    // L(X) = (sum_s v^. z_s (V_s - r_s(u)) - Z_T(u) h(X)) / z_0, then / (X - u)
    PCHECK(hipMemsetAsync(fz, 0, nb, st));
    for (int set = 0; set < nsets; set++) {
        PCHECK(zk_fr_op(V[set], V[set], ocs(32 + 9 + 2 * set), OP_SUB, 1, st));
        PCHECK(field_op(fz, fz, 0, j.cf(J_V), V[set], 0, ocs(32 + 10 + 2 * set), n, FO_ADD, st));
    }
    PCHECK(field_op(fz, fz, 0, nullptr, hx, 0, ocs(32 + 15), n, FO_SUB, st));
    PCHECK(field_op(fz, fz, 0, ocs(32 + 16), nullptr, 0, nullptr, n, FO_LEFT, st));
    PCHECK(zk_kate_divide(fz, fz, j.cf(J_U), ocs(0), n, kate_tmp, st));
    PCHECK(commit(j, pk.g, std::vector<const void *>{fz}, out + (size_t)ncom * 64, batched, st));
    tr.points(out + (size_t)ncom * 64, 1);
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

// Commitments a proof writes at most: advice, m, logup z, permutation z,
// shuffle z, quotient pieces, then the opening (GWC: one per distinct point;
// Shplonk: h and L).
int zkp_commitment_count(const zkp_circuit *c) {
    Shape s;
    if (!make_shape(c, s)) return 0;
    const int open = open_gwc_points(s) > 2 ? open_gwc_points(s) : 2;
    return s.A + 2 * s.L + s.Sp + s.S + kQuot + open;
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
    hipError_t e = hipFree(j->mem);
    delete j;
    return e;
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
        // reference writes them; every challenge is squeezed from the transcript
        Transcript tr;

        // ---- 0. witness (the caller's advice in the reference) -----------------
        PCHECK(witness(j, st));
        if (flags & ZKP_TAMPER_GATE) {
            // self-test of the quotient check: one usable row breaks the gate
            uint8_t *d = (uint8_t *)j.adv_lag[s.col_gate + 3] + (size_t)(s.u / 2) * FR;
            PCHECK(zk_fr_op(d, d, pk.cf(C_ONE), OP_ADD, 1, st));
        }
        PCHECK(mark(0));
        const double t_start = tp;

        // ---- 1. advice commitments (Lagrange basis) and iNTT --------------------
        {
            std::vector<const void *> polys(j.adv_lag.begin(), j.adv_lag.end());
            PCHECK(commit(j, pk.g_lag, polys, out + (size_t)ncom * 64, batched != 0, st));
            tr.points(out + (size_t)ncom * 64, s.A);
            ncom += s.A;
            (void)tr.challenge();  // theta (src/lib.rs:395): this circuit's lookups are single columns
            PCHECK(intt_all(j, j.adv_lag, j.adv_coef, j.adv_tmp, st));
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
            tr.points(out + (size_t)ncom * 64, s.L);
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
        for (int l = 0; l < s.L; l++) {
            PCHECK(logup_sum_input_inv(j.w_sum, j.adv_lag[s.col_lookup + l], j.w_temp, beta, 0, (int)s.n, st));
            PCHECK(hipMemcpyAsync(j.w_tab, pk.t_lag, (size_t)s.n * FR, hipMemcpyDeviceToDevice, st));
            PCHECK(hipMemsetAsync(j.cf(J_LASTZ), 0, FR, st));
            PCHECK(eval_logup_z_compose(j.zl_lag[l], j.w_sum, j.w_tab, j.m_lag[l], beta, j.cf(J_LASTZ), (int)s.u,
                                        (int)s.n, st));
        }
        // device pointer arrays: [values | sigmas | zs], then the gate term list,
        // its rotations, the permutation sets, the permutation columns' cosets and
        // the shuffle inputs' / tables' cosets; shuffle z inputs reuse the start
        void **pa = j.ptrs;
        void **pa_gate = pa + 2 * s.P + s.Sp, **pa_sets = pa_gate + 16 + 8;
        void **pa_pv = pa_sets + s.Sp, **pa_ps = pa_pv + s.P;
        void **pa_sh = pa_ps + s.P;
        int *pa_rots = (int *)(pa_gate + 16);
        if (s.P > 0) {
            std::vector<void *> host(2 * s.P + s.Sp);
            for (int c = 0; c < s.P; c++) {
                host[c] = j.adv_lag[s.col_perm + c];
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
            // permutation, lookup, shuffle: the reference's transcript order
            // (src/lib.rs:968-980)
            std::vector<const void *> polys;
            for (auto p : j.zp_lag) polys.push_back(p);
            for (auto p : j.zl_lag) polys.push_back(p);
            for (auto p : j.zs_lag) polys.push_back(p);
            if (!polys.empty()) PCHECK(commit(j, pk.g_lag, polys, out + (size_t)ncom * 64, batched != 0, st));
            tr.points(out + (size_t)ncom * 64, (int)polys.size());
            ncom += (int)polys.size();
            const zkh::Fr yc = tr.challenge();  // y (src/lib.rs:988)
            PCHECK(put_challenges(j, J_Y, &yc, 1, st));
            memcpy(inf.challenges[2], &yc, FR);
            // closure of every grand product at row u: logup 0, permutation (last
            // set) 1, shuffle 1
            int bad = 0;  // bit 0: a logup sum, bit 1: the permutation, bit 2: a shuffle
            uint64_t one[4], got[4];
            PCHECK(hipMemcpyAsync(one, pk.cf(C_ONE), FR, hipMemcpyDeviceToHost, st));
            PCHECK(hipStreamSynchronize(st));
            auto at_u = [&](const void *z) -> hipError_t {
                PCHECK(hipMemcpyAsync(got, (const uint8_t *)z + (size_t)s.u * FR, FR, hipMemcpyDeviceToHost, st));
                return hipStreamSynchronize(st);
            };
            for (auto z : j.zl_lag) {
                PCHECK(at_u(z));
                if (got[0] | got[1] | got[2] | got[3]) bad |= 1;
            }
            if (s.Sp) {
                PCHECK(at_u(j.zp_lag[s.Sp - 1]));
                if (memcmp(got, one, FR) != 0) bad |= 2;
            }
            for (auto z : j.zs_lag) {
                PCHECK(at_u(z));
                if (memcmp(got, one, FR) != 0) bad |= 4;
            }
            inf.z_ok = bad ? 0 : 1;
            inf.z_fail = bad;
            PCHECK(intt_all(j, j.zl_lag, j.zl_coef, j.zl_tmp, st));
            PCHECK(intt_all(j, j.zp_lag, j.zp_coef, j.zp_tmp, st));
            PCHECK(intt_all(j, j.zs_lag, j.zs_coef, j.zs_tmp, st));
        }
        PCHECK(mark(3));

        // ---- 4. h(X) on the extended coset ----------------------------------------
        // ext slots: gate a,b,c,d | f_l | m_l | zl_l | perm v | zp | shuffle in, sh | zs
        const int E_GATE = 0, E_F = 4, E_M = E_F + s.L, E_ZL = E_M + s.L, E_V = E_ZL + s.L, E_ZP = E_V + s.P,
                  E_IN = E_ZP + s.Sp, E_ZS = E_IN + 2 * s.S;
        {
            std::vector<const void *> c;
            for (int i = 0; i < 4; i++) c.push_back(j.adv_coef[s.col_gate + i]);
            for (int l = 0; l < s.L; l++) c.push_back(j.adv_coef[s.col_lookup + l]);
            for (auto p : j.m_coef) c.push_back(p);
            for (auto p : j.zl_coef) c.push_back(p);
            for (int i = 0; i < s.P; i++) c.push_back(j.adv_coef[s.col_perm + i]);
            for (auto p : j.zp_coef) c.push_back(p);
            for (int t = 0; t < 2 * s.S; t++) c.push_back(j.adv_coef[s.col_shuffle + t]);
            for (auto p : j.zs_coef) c.push_back(p);
            PCHECK(extend_all(j, c, 0, st));
        }
        void *y = j.cf(J_Y);
        const int rot1 = 1 << kExtBits;  // w X on the coset
        const int N = (int)s.N;
        PCHECK(hipMemsetAsync(j.h, 0, (size_t)s.N * FR, st));
        // gate: q a b + q c - q d
        {
            void *hv[13] = {pk.cf(C_ONE), pk.q_ext, j.ext[E_GATE], j.ext[E_GATE + 1], nullptr,
                            pk.cf(C_ONE), pk.q_ext, j.ext[E_GATE + 2], nullptr,
                            pk.cf(C_MONE), pk.q_ext, j.ext[E_GATE + 3], nullptr};
            int rots[7] = {0, 0, 0, 0, 0, 0, 0};
            PCHECK(hipMemcpyAsync(pa_gate, hv, sizeof(hv), hipMemcpyHostToDevice, st));
            PCHECK(hipMemcpyAsync(pa_rots, rots, sizeof(rots), hipMemcpyHostToDevice, st));
            PCHECK(field_op_batch_mul_sum(j.h, pa_gate, pa_rots, 13, N, st));
            PCHECK(hipStreamSynchronize(st));
        }
        // permutation
        if (s.Sp > 0) {
            PCHECK(zk_permutation_eval_h_p1_on(j.h, j.ext[E_ZP], j.ext[E_ZP + s.Sp - 1], pk.l0_ext, pk.llast_ext, y, N,
                                               st));
            std::vector<void *> sets(s.Sp);
            for (int t = 0; t < s.Sp; t++) sets[t] = j.ext[E_ZP + t];
            PCHECK(hipMemcpyAsync(pa_sets, sets.data(), (size_t)s.Sp * 8, hipMemcpyHostToDevice, st));
            PCHECK(zk_permutation_eval_h_p2_on(j.h, pa_sets, pk.l0_ext, pk.llast_ext, y, s.Sp, (int)s.u * rot1, N, st));
            PCHECK(hipStreamSynchronize(st));
            // per set: the fused product term (sigma and X resident as cosets)
            std::vector<void *> cols(2 * s.P);
            for (int c = 0; c < s.P; c++) {
                cols[c] = j.ext[E_V + c];
                cols[s.P + c] = pk.sig_ext[c];
                PCHECK(zk_fr_op(j.cf(J_BDELTA + c), beta, pk.cf(C_DELTAS + c), OP_MUL, 1, st));
            }
            PCHECK(hipMemcpyAsync(pa_pv, cols.data(), cols.size() * 8, hipMemcpyHostToDevice, st));
            for (int t = 0; t < s.Sp; t++) {
                int c0 = t * kChunk, nc = s.P - c0 < kChunk ? s.P - c0 : kChunk;
                PCHECK(zk_permutation_eval_h_set(j.h, j.ext[E_ZP + t], pa_pv + c0, pa_ps + c0, nc, beta, gamma,
                                                 j.cf(J_BDELTA + c0), pk.x_ext, pk.lact_ext, y, rot1, N, st));
            }
            PCHECK(hipStreamSynchronize(st));  // the host pointer array
        }
        // logup: (t + beta) once (every lookup reads the one table), (f + beta),
        // product-sum 1
        if (s.L > 0) PCHECK(field_op(j.aux1, pk.t_ext, 0, nullptr, nullptr, 0, beta, N, 0, st));
        for (int l = 0; l < s.L; l++) {
            PCHECK(field_op(j.aux0, j.ext[E_F + l], 0, nullptr, nullptr, 0, beta, N, 0, st));
            PCHECK(logup_eval_h(j.h, j.aux0, pk.ones_ext, j.aux1, j.ext[E_M + l], j.ext[E_ZL + l], j.ext[E_ZL + l],
                                pk.l0_ext, pk.llast_ext, pk.lact_ext, y, rot1, N, st));
        }
        // shuffle: each group one plain input column and one plain table column,
        // the case the reference hands to shuffle_eval_h_v2 with betas [beta, ...]
        // (src/eval_h.rs:1095-1152); the kernel adds beta itself
        if (s.S > 0) {
            std::vector<void *> sh(2 * s.S);
            for (int t = 0; t < s.S; t++) {
                sh[t] = j.ext[E_IN + 2 * t];
                sh[s.S + t] = j.ext[E_IN + 2 * t + 1];
            }
            PCHECK(hipMemcpyAsync(pa_sh, sh.data(), sh.size() * 8, hipMemcpyHostToDevice, st));
            for (int t = 0; t < s.S; t++)
                PCHECK(zk_shuffle_eval_h_v2_on(j.h, pa_sh + t, pa_sh + s.S + t, beta, 1, j.ext[E_ZS + t], pk.l0_ext,
                                               pk.llast_ext, pk.lact_ext, y, rot1, N, st));
            PCHECK(hipStreamSynchronize(st));  // the host pointer array
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
            PCHECK(hipMemsetAsync(j.flag, 0, 4, st));
            PCHECK(zk::wit_any_nonzero(c, (uint64_t)kQuot * s.n, s.N, j.flag, st));
            uint32_t flag = 1;
            PCHECK(hipMemcpyAsync(&flag, j.flag, 4, hipMemcpyDeviceToHost, st));
            PCHECK(hipStreamSynchronize(st));
            inf.quotient_ok = flag == 0 ? 1 : 0;
            std::vector<const void *> pieces;
            for (int i = 0; i < kQuot; i++) pieces.push_back((uint8_t *)c + (size_t)i * s.n * FR);
            PCHECK(commit(j, pk.g, pieces, out + (size_t)ncom * 64, batched != 0, st));
            tr.points(out + (size_t)ncom * 64, kQuot);
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
