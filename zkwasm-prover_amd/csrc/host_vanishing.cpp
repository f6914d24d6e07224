// Host-side mirror of the vanishing construction (src/eval_h.rs:464-583):
// the reference's call sequence over this library's C-ABI.
//   1. h *= t_evaluations[i mod t_len]          (1/Z_H on the extended coset)
//   2. extended iNTT with the extended divisor   (ntt, n_inv set)
//   3. coset un-distribution [g_coset_inv, g_coset] over the whole extended
//      domain (extended_prepare, to_coset = 1)
//   4. one commitment per quotient piece of n coefficients against g
//   5. h_pieces = sum_i piece_i * xn^i by Horner (field_op add with l_c = xn)
// Steps 1-3 run on the stream the reference uses (the legacy default
// stream); step 4 runs either the reference driver's per-piece MSMs or one
// batched pass (zk_msm_batch).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "guard.h"
#include "zkwasm_prover_kernel.h"

namespace {

#define VCHECK(x)                             \
    do {                                      \
        hipError_t e_ = (x);                  \
        if (e_ != hipSuccess) return e_;      \
    } while (0)

struct Scratch {
    void *p = nullptr;
    ~Scratch() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

extern "C" {

// Steps 1-4.  h (2^log_ext Fr) and tmp (same size) are device buffers; on
// return *coeffs is whichever of the two holds the coset coefficients, whose
// first quotient_degree * 2^log_n entries are the pieces.  commitments (host)
// receives quotient_degree affine points (64 B each, identity = (0, 0)).
hipError_t zkp_vanishing_commit(void *h, void *tmp, const void *t_evals, int t_len, const void *ext_intt_omegas,
                                const void *ext_divisor, const void *coset_powers, int log_n, int log_ext,
                                int quotient_degree, const void *g, void *commitments, void **coeffs, int batched) {
    return zk::guard([&]() -> hipError_t {
        if (!h || !tmp || !t_evals || t_len <= 0 || !ext_intt_omegas || !ext_divisor || !coset_powers || !g ||
            !commitments || !coeffs || log_n < 1 || log_ext < log_n || log_ext > 30 || quotient_degree < 1 ||
            ((int64_t)quotient_degree << log_n) > ((int64_t)1 << log_ext))
            return hipErrorInvalidValue;
        const int n = 1 << log_n, N = 1 << log_ext;
        // eval_h.rs:467-481
        VCHECK(field_mul_zip(h, t_evals, t_len, N));
        // eval_h.rs:484-505
        bool swap = false;
        VCHECK(ntt(h, tmp, nullptr, ext_intt_omegas, ext_divisor, log_ext, 8, &swap, nullptr));
        void *c = swap ? tmp : h;
        // eval_h.rs:507-510
        VCHECK(extended_prepare(c, coset_powers, 3, n, N, 1, nullptr));
        VCHECK(hipStreamSynchronize(nullptr));
        *coeffs = c;
        // eval_h.rs:512-540
        const size_t piece = (size_t)n * 32;
        if (!batched) {
            std::vector<const void *> sets(quotient_degree);
            for (int i = 0; i < quotient_degree; i++) sets[i] = (const uint8_t *)c + i * piece;
            return zkp_batch_msm(commitments, g, sets.data(), quotient_degree, n, 0);
        }
        std::vector<zk_msm_item> items(quotient_degree);
        for (int i = 0; i < quotient_degree; i++) items[i] = zk_msm_item{g, (const uint8_t *)c + i * piece, n, 0};
        size_t bytes = zk_msm_batch_scratch_bytes(items.data(), quotient_degree, 0);
        Scratch scratch, out;
        VCHECK(hipMalloc(&scratch.p, bytes));
        VCHECK(hipMalloc(&out.p, (size_t)quotient_degree * ZK_MSM_OUT_BYTES));
        VCHECK(zk_msm_batch(out.p, items.data(), quotient_degree, 0, scratch.p, bytes, nullptr));
        std::vector<uint8_t> jac(96);
        for (int i = 0; i < quotient_degree; i++) {
            VCHECK(hipMemcpy(jac.data(), (uint8_t *)out.p + (size_t)i * ZK_MSM_OUT_BYTES, 96, hipMemcpyDeviceToHost));
            uint8_t *o = (uint8_t *)commitments + (size_t)i * 64;
            bool zero_z = true;
            for (int k = 64; k < 96; k++) zero_z &= jac[k] == 0;
            if (zero_z)
                memset(o, 0, 64);
            else
                memcpy(o, jac.data(), 64);
        }
        return hipSuccess;
    });
}

// Step 5 (eval_h.rs:546-583): folds the pieces of coeffs into the last piece,
// last = last * xn + piece_i for i = quotient_degree - 2 .. 0, and copies the
// n results to h_pieces (host).  xn is a device pointer to x^n.
hipError_t zkp_vanishing_pieces(void *coeffs, const void *xn, int log_n, int quotient_degree, void *h_pieces) {
    return zk::guard([&]() -> hipError_t {
        if (!coeffs || !xn || !h_pieces || log_n < 1 || log_n > 30 || quotient_degree < 1) return hipErrorInvalidValue;
        const int n = 1 << log_n;
        const size_t piece = (size_t)n * 32;
        uint8_t *last = (uint8_t *)coeffs + (size_t)(quotient_degree - 1) * piece;
        for (int i = quotient_degree - 2; i >= 0; i--)
            VCHECK(field_op(last, last, 0, xn, (uint8_t *)coeffs + (size_t)i * piece, 0, nullptr, n, 0, nullptr));
        return hipMemcpy(h_pieces, last, piece, hipMemcpyDeviceToHost);
    });
}

}  // extern "C"
