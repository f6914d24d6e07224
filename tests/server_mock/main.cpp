// Scenarios: 8 proofs of two shapes one at a time, then all at once on 3
// slots with a budget for about two proof states, submitting while workers
// run (eager admission); the same through the cross-proof batcher; a lone
// request never gets company at light load; a large request that does not
// fit is passed at most 8 times before it starts; states pooled at creation
// are reused with none created or freed while serving.  Exits non-zero on any
// error; the commitments of each seed must match.
#include <hip/hip_runtime.h>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <mutex>
#include <thread>
#include <vector>
#include "zkwasm_prover_kernel.h"

extern std::mutex g_log_mu;
extern std::vector<unsigned long long> g_start_log;
extern std::atomic<int> g_hook_errors;
extern std::atomic<int> g_job_creates, g_job_destroys;

static int mix(int flags) {
    zkp_circuit a{10, 30, 2, 1, 3, 7, 6, 2, 7, 5}, b{11, 30, 2, 1, 3, 8, 6, 2, 8, 5};
    zkp_circuit shapes[2] = {a, b};
    int plan[8][2] = {{0, 101}, {1, 102}, {0, 103}, {1, 104}, {1, 105}, {0, 106}, {0, 107}, {1, 108}};
    void *s1;
    if (zks_create(shapes, 2, 1, 0, flags, 1, &s1)) return 1;
    for (auto &p : plan) {
        std::vector<uint8_t> com(64 * 32);
        zks_result r;
        int id = zks_submit(s1, p[0], p[1]);
        if (zks_wait(s1, id, com.data(), &r)) return 2;
    }
    (void)zks_destroy(s1);
    size_t jb = zkp_job_bytes(&b);
    void *s3;
    if (zks_create(shapes, 2, 3, 2 * jb + jb / 2, flags, 1, &s3)) return 3;
    int ids[8];
    for (int i = 0; i < 8; i++) ids[i] = zks_submit(s3, plan[i][0], plan[i][1]);
    for (int i = 0; i < 8; i++) {
        std::vector<uint8_t> com(64 * 32);
        zks_result r;
        if (zks_wait(s3, ids[i], com.data(), &r)) return 4;
        if (com[0] != (uint8_t)(plan[i][1] & 0xff)) return 5;
    }
    int pr; size_t pb, bud;
    (void)zks_stats(s3, &pr, &pb, &bud);
    if (pr < 2 || pb > bud) return 6;
    unsigned long long launches, merged, items;
    (void)zks_batch_stats(s3, &launches, &merged, &items);
    printf("flags %x peak running %d bytes %zu budget %zu batcher %llu/%llu/%llu\n", flags, pr, pb, bud, launches,
           merged, items);
    if ((flags & ZKS_CROSS_BATCH) && items != 8 * 6) return 7;
    (void)zks_destroy(s3);
    return g_hook_errors.load() ? 8 : 0;
}

// light load: one request at a time with gaps; none may overlap another
static int alone_at_light_load() {
    zkp_circuit a{10, 30, 2, 1, 3, 7, 6, 2, 7, 5};
    void *s;
    if (zks_create(&a, 1, 3, 0, ZKP_BATCHED, 1, &s)) return 11;
    int first = zks_submit(s, 0, 300);
    std::this_thread::sleep_for(std::chrono::milliseconds(1));
    int second = zks_submit(s, 0, 301);  // one waiting behind one running: no overlap
    zks_result r1, r2;
    if (zks_wait(s, first, nullptr, &r1) || zks_wait(s, second, nullptr, &r2)) return 12;
    int pr; size_t pb, bud;
    (void)zks_stats(s, &pr, &pb, &bud);
    (void)zks_destroy(s);
    printf("light load: peak running %d\n", pr);
    return pr == 1 ? 0 : 13;
}

// a big request that never fits beside the small ones in flight starts after
// at most 8 later requests have passed it
static int no_starvation() {
    zkp_circuit a{10, 30, 2, 1, 3, 7, 6, 2, 7, 5}, b{12, 30, 2, 1, 3, 8, 6, 2, 8, 5};
    zkp_circuit shapes[2] = {a, b};
    void *s;
    size_t big = zkp_job_bytes(&b);  // 4 small states
    if (zks_create(shapes, 2, 3, big, ZKP_BATCHED | ZKS_EAGER, 1, &s)) return 21;
    {
        std::lock_guard<std::mutex> lk(g_log_mu);
        g_start_log.clear();
    }
    std::vector<int> ids;
    ids.push_back(zks_submit(s, 0, 500));
    ids.push_back(zks_submit(s, 1, 999));
    for (int i = 0; i < 24; i++) ids.push_back(zks_submit(s, 0, 501 + i));
    for (int id : ids)
        if (zks_wait(s, id, nullptr, nullptr)) return 22;
    (void)zks_destroy(s);
    size_t pos = 0;
    std::lock_guard<std::mutex> lk(g_log_mu);
    while (pos < g_start_log.size() && g_start_log[pos] != 999) pos++;
    printf("big request started %zu-th of %zu\n", pos, g_start_log.size());
    return pos <= 1 + 8 ? 0 : 23;
}

// states are created with the server (one per shape, then up to `slots` per
// shape, smallest first, within the budget): a burst of mixed requests then
// creates and frees none, and each shape runs at most its state count at once
static int pooled_states() {
    zkp_circuit a{10, 30, 2, 1, 3, 7, 6, 2, 7, 5}, b{12, 30, 2, 1, 3, 8, 6, 2, 8, 5};
    zkp_circuit shapes[2] = {a, b};
    const size_t ja = zkp_job_bytes(&a), jb = zkp_job_bytes(&b);
    void *s;
    const int c0 = g_job_creates.load();
    if (zks_create(shapes, 2, 3, jb + 3 * ja, ZKP_BATCHED | ZKS_EAGER, 1, &s)) return 31;
    const int c1 = g_job_creates.load(), d1 = g_job_destroys.load();
    if (c1 - c0 != 4) return 32;  // b once, a three times
    std::vector<int> ids;
    for (int i = 0; i < 12; i++) ids.push_back(zks_submit(s, i % 3 == 0 ? 1 : 0, 700 + i));
    for (int id : ids)
        if (zks_wait(s, id, nullptr, nullptr)) return 33;
    int pr; size_t pb, bud;
    (void)zks_stats(s, &pr, &pb, &bud);
    const int c2 = g_job_creates.load(), d2 = g_job_destroys.load();
    (void)zks_destroy(s);
    printf("pooled states: created %d while serving, freed %d, peak running %d, bytes %zu of %zu\n", c2 - c1,
           d2 - d1, pr, pb, bud);
    return (c2 == c1 && d2 == d1 && pr >= 2 && pb <= bud) ? 0 : 34;
}

int main() {
    int r = mix(ZKP_BATCHED | ZKS_EAGER);
    if (!r) r = mix(ZKP_BATCHED | ZKS_EAGER | ZKS_CROSS_BATCH);
    if (!r) r = alone_at_light_load();
    if (!r) r = no_starvation();
    if (!r) r = pooled_states();
    if (r) printf("failed: %d\n", r);
    return r;
}
