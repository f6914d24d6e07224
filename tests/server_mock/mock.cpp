// CPU stand-ins for the HIP runtime calls and the prover core that the proof
// server uses, so its scheduling logic (admission, pooling, concurrent
// submits while workers run) can be exercised under the sanitizers without a
// GPU.  Proofs sleep a few ms and write seed-derived bytes.
#include <hip/hip_runtime.h>
#include <chrono>
#include <thread>
#include <cstring>
#include <atomic>
#include "zkwasm_prover_kernel.h"
extern "C" {
hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipMemGetInfo(size_t *f, size_t *t) { *f = (size_t)100 << 30; *t = (size_t)288 << 30; return hipSuccess; }
hipError_t hipStreamCreate(hipStream_t *s) { *s = (hipStream_t)new int(0); return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t s) { delete (int *)s; return hipSuccess; }
size_t zkp_job_bytes(const zkp_circuit *c) { return (size_t)1 << (c->k + 10); }
int zkp_commitment_count(const zkp_circuit *c) { return 5 + c->k; }
hipError_t zkp_pk_create(const zkp_circuit *, unsigned long long, hipStream_t, void **pk) { *pk = new int(1); return hipSuccess; }
hipError_t zkp_pk_destroy(void *pk) { delete (int *)pk; return hipSuccess; }
std::atomic<int> live{0};
hipError_t zkp_job_create(void *, void **job) { *job = new int(2); live++; return hipSuccess; }
hipError_t zkp_job_destroy(void *job) { delete (int *)job; live--; return hipSuccess; }
hipError_t zkp_prove(void *job, unsigned long long seed, int, void *com, void *, zkp_proof_info *info, hipStream_t) {
    std::this_thread::sleep_for(std::chrono::milliseconds(5 + seed % 7));
    memset(com, (int)(seed & 0xff), 64);
    memset(info, 0, sizeof(*info)); info->z_ok = info->quotient_ok = 1; info->ncommit = 1;
    return hipSuccess;
}
}
