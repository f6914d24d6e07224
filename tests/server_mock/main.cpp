// Scenario: 8 proofs of two shapes one at a time, then all at once on 3 slots
// with a budget for about two proof states, submitting while workers run.
// Exits non-zero on any error; the commitments of each seed must match.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <vector>
#include "zkwasm_prover_kernel.h"
int main() {
    zkp_circuit a{10, 30, 2, 1, 3, 7, 6, 2, 7, 5}, b{11, 30, 2, 1, 3, 8, 6, 2, 8, 5};
    zkp_circuit shapes[2] = {a, b};
    int plan[8][2] = {{0, 101}, {1, 102}, {0, 103}, {1, 104}, {1, 105}, {0, 106}, {0, 107}, {1, 108}};
    void *s1;
    if (zks_create(shapes, 2, 1, 0, 1, 1, &s1)) return 1;
    for (auto &p : plan) {
        std::vector<uint8_t> com(64 * 32);
        zks_result r;
        int id = zks_submit(s1, p[0], p[1]);
        if (zks_wait(s1, id, com.data(), &r)) return 2;
        printf("alone %d ok lat %.1f\n", p[1], r.latency_ms);
    }
    (void)zks_destroy(s1);
    size_t jb = zkp_job_bytes(&b);
    void *s3;
    if (zks_create(shapes, 2, 3, 2 * jb + jb / 2, 1, 1, &s3)) return 3;
    int ids[8];
    for (int i = 0; i < 8; i++) ids[i] = zks_submit(s3, plan[i][0], plan[i][1]);
    for (int i = 0; i < 8; i++) {
        std::vector<uint8_t> com(64 * 32);
        zks_result r;
        if (zks_wait(s3, ids[i], com.data(), &r)) return 4;
        if (com[0] != (uint8_t)(plan[i][1] & 0xff)) return 5;
        printf("mix %d ok lat %.1f\n", plan[i][1], r.latency_ms);
    }
    int pr; size_t pb, bud;
    (void)zks_stats(s3, &pr, &pb, &bud);
    if (pr < 2 || pb > bud) return 6;
    printf("peak running %d bytes %zu budget %zu\n", pr, pb, bud);
    (void)zks_destroy(s3);
    return 0;
}
