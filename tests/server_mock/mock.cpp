// CPU stand-ins for the HIP runtime calls and the prover core that the proof
// server uses, so its scheduling logic (admission, pooling, concurrent
// submits while workers run, the cross-proof MSM batcher) can be exercised
// under the sanitizers without a GPU.  Proofs sleep a few ms and write
// seed-derived bytes; a job with a commit hook sends two groups of fake MSM
// items through it and checks that each result comes back to its own items.
#include <hip/hip_runtime.h>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>
#include "zkinternal.h"
#include "zkwasm_prover_kernel.h"

std::mutex g_log_mu;
std::vector<unsigned long long> g_start_log;  // seeds in the order their proofs started
std::atomic<int> g_hook_errors{0};

struct MockJob {
    zk::CommitFn fn = nullptr;
    void *ctx = nullptr;
};

namespace zk {
hipError_t job_set_commit_fn(void *job, CommitFn fn, void *ctx) {
    ((MockJob *)job)->fn = fn;
    ((MockJob *)job)->ctx = ctx;
    return hipSuccess;
}
}  // namespace zk

extern "C" {
hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipMemGetInfo(size_t *f, size_t *t) { *f = (size_t)100 << 30; *t = (size_t)288 << 30; return hipSuccess; }
hipError_t hipStreamCreate(hipStream_t *s) { *s = (hipStream_t)new int(0); return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t s) { delete (int *)s; return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned int) { return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { *e = (hipEvent_t)new int(0); return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventDestroy(hipEvent_t e) { delete (int *)e; return hipSuccess; }
hipError_t hipMalloc(void **p, size_t n) { *p = malloc(n ? n : 1); return *p ? hipSuccess : hipErrorOutOfMemory; }
hipError_t hipFree(void *p) { free(p); return hipSuccess; }
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t) {
    memcpy(d, s, n);
    return hipSuccess;
}
size_t zk_msm_batch_scratch_bytes(const zk_msm_item *items, int count, int) {
    size_t b = 0;
    for (int i = 0; i < count; i++) b += (size_t)items[i].n;
    return b;
}
int zk_msm_table_window(int) { return 0; }  // no fixed-base tables in the mock
// result m: x = the item's scalar pointer repeated, z = 1 (not the identity)
hipError_t zk_msm_batch(void *out, const zk_msm_item *items, int count, int, void *, size_t scratch_bytes,
                        hipStream_t) {
    if (zk_msm_batch_scratch_bytes(items, count, 0) > scratch_bytes) return hipErrorInvalidValue;
    for (int m = 0; m < count; m++) {
        uint8_t *o = (uint8_t *)out + (size_t)m * ZK_MSM_OUT_BYTES;
        memset(o, 0, 96);
        for (int w = 0; w < 8; w++) memcpy(o + 8 * w, &items[m].scalars, 8);
        o[64] = 1;
    }
    return hipSuccess;
}
size_t zkp_job_bytes(const zkp_circuit *c) { return (size_t)1 << (c->k + 10); }
int zkp_commitment_count(const zkp_circuit *c) { return 5 + c->k; }
hipError_t zkp_pk_create(const zkp_circuit *, unsigned long long, hipStream_t, void **pk) { *pk = new int(1); return hipSuccess; }
hipError_t zkp_pk_destroy(void *pk) { delete (int *)pk; return hipSuccess; }
std::atomic<int> live{0};
std::atomic<int> g_job_creates{0}, g_job_destroys{0};
hipError_t zkp_job_create(void *, void **job) {
    *job = new MockJob();
    live++;
    g_job_creates++;
    return hipSuccess;
}
hipError_t zkp_job_destroy(void *job) {
    delete (MockJob *)job;
    live--;
    g_job_destroys++;
    return hipSuccess;
}
hipError_t zkp_prove(void *job, unsigned long long seed, int, void *com, void *, zkp_proof_info *info,
                     hipStream_t st) {
    {
        std::lock_guard<std::mutex> lk(g_log_mu);
        g_start_log.push_back(seed);
    }
    MockJob *j = (MockJob *)job;
    if (j->fn) {
        for (int g = 0; g < 2; g++) {
            zk_msm_item it[3];
            uint8_t out[3 * 64];
            for (int m = 0; m < 3; m++)
                it[m] = zk_msm_item{nullptr, (const void *)(uintptr_t)(seed * 16 + g * 4 + m + 1), 1 << 10, 0};
            if (j->fn(j->ctx, it, 3, out, st) != hipSuccess) g_hook_errors++;
            for (int m = 0; m < 3; m++)
                if (memcmp(out + m * 64, &it[m].scalars, 8) != 0) g_hook_errors++;
        }
    }
    std::this_thread::sleep_for(std::chrono::milliseconds(5 + seed % 7));
    memset(com, (int)(seed & 0xff), 64);
    memset(info, 0, sizeof(*info)); info->z_ok = info->quotient_ok = 1; info->ncommit = 1;
    return hipSuccess;
}
}
