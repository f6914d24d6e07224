// Host residency: a pinned host pool and an asynchronous device-to-host copy
// queue (the reference's src/hugetlb.rs and src/async_copy_queue.rs).
//
// Pool: process-wide, slabs from hipHostMalloc (page-locked and mapped for
// DMA at allocation), split as host_pool.h describes.  Host columns the
// prover reads from or writes to (witness upload, h pieces, lookup columns)
// placed here move at DMA rate and let hipMemcpyAsync return at once.
//
// Copy queue: its own stream (default priority).  A push orders the copy
// after everything enqueued so far on the producer's stream through an event
// (the host never waits for the producer), then copies on the queue's
// stream, so the producer's next kernels overlap the transfer.  At most
// max_inflight copies are outstanding; a push into a full queue waits for the
// oldest (the only host wait).  Modes:
//   default   the caller leaves dev_src untouched until sync_all;
//   SNAPSHOT  the queue first copies dev_src into a staging buffer of its own
//             and makes the producer's stream wait for that device copy only,
//             so later work on the producer's stream may overwrite dev_src
//             (push_with_copy_buffer, async_copy_queue.rs:42-55);
//   RELEASE   the queue owns dev_src (from hipMalloc) and frees it once the
//             copy has landed (push / push_with_new_buffer, :57-82).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <chrono>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "guard.h"
#include "host_pool.h"
#include "zkwasm_prover_kernel.h"

namespace {

struct HipHost {
    static void *alloc(size_t bytes) {
        void *p = nullptr;
        return hipHostMalloc(&p, bytes, hipHostMallocDefault) == hipSuccess ? p : nullptr;
    }
    static void release(void *p) { (void)hipHostFree(p); }
};

// never destroyed: pinned pages are returned with the process, after the
// runtime's own teardown would have made hipHostFree unsafe
zkh::HostPool<HipHost> &pool() {
    static auto *p = new zkh::HostPool<HipHost>();
    return *p;
}

struct Stage {
    void *p = nullptr;
    size_t cap = 0;
    bool busy = false;
};

struct Entry {
    hipEvent_t ready = nullptr, snap = nullptr, done = nullptr;
    void *release = nullptr;
    int stage = -1;
};

struct CopyQueue {
    hipStream_t cs = nullptr;
    int max = 0;
    std::deque<Entry> q;
    std::vector<Stage> stages;
    unsigned long long pushed = 0, full_waits = 0;
    double max_wait_ms = 0;
    std::mutex mu;
};

hipError_t finish(CopyQueue &cq, Entry &e) {
    hipError_t err = hipEventSynchronize(e.done);
    if (e.release) {
        hipError_t f = hipFree(e.release);
        if (err == hipSuccess) err = f;
    }
    if (e.stage >= 0) cq.stages[e.stage].busy = false;
    for (hipEvent_t ev : {e.ready, e.snap, e.done})
        if (ev) (void)hipEventDestroy(ev);
    return err;
}

hipError_t drain(CopyQueue &cq) {
    hipError_t first = hipSuccess;
    while (!cq.q.empty()) {
        Entry e = cq.q.front();
        cq.q.pop_front();
        hipError_t err = finish(cq, e);
        if (first == hipSuccess) first = err;
    }
    return first;
}

int take_stage(CopyQueue &cq, size_t bytes, hipError_t *err) {
    int pick = -1;
    for (size_t i = 0; i < cq.stages.size(); i++) {
        Stage &s = cq.stages[i];
        if (s.busy) continue;
        if (s.cap >= bytes) {
            pick = (int)i;
            break;
        }
        if (pick < 0) pick = (int)i;  // a free one to grow
    }
    if (pick < 0) {
        try {
            cq.stages.push_back(Stage{});
        } catch (...) {
            *err = hipErrorOutOfMemory;
            return -1;
        }
        pick = (int)cq.stages.size() - 1;
    }
    Stage &s = cq.stages[pick];
    if (s.cap < bytes) {
        if (s.p) (void)hipFree(s.p);
        s.p = nullptr;
        s.cap = 0;
        *err = hipMalloc(&s.p, bytes);
        if (*err != hipSuccess) return -1;
        s.cap = bytes;
    }
    s.busy = true;
    return pick;
}

}  // namespace

extern "C" {

hipError_t zk_pinned_reserve(size_t slabs) {
    return zk::guard([&]() -> hipError_t {
        return pool().reserve(slabs) ? hipSuccess : hipErrorOutOfMemory;
    });
}

hipError_t zk_pinned_alloc(size_t bytes, void **out) {
    return zk::guard([&]() -> hipError_t {
        if (!out || bytes == 0 || bytes > zkh::kSlabBytes) return hipErrorInvalidValue;
        *out = pool().alloc(bytes);
        return *out ? hipSuccess : hipErrorOutOfMemory;
    });
}

hipError_t zk_pinned_free(void *p) { return pool().release(p) ? hipSuccess : hipErrorInvalidValue; }

// src/lib.rs:107-112 / src/buffer.rs:19-34: `count` zeroed columns of `rows`
// Fr (32 B) from the pinned pool, zeroed by up to 16 threads as the
// reference's rayon loop does; on failure nothing stays allocated.
hipError_t zk_prepare_advice_buffer(int rows, int count, void **cols) {
    return zk::guard([&]() -> hipError_t {
        if (rows <= 0 || count <= 0 || !cols) return hipErrorInvalidValue;
        const size_t bytes = (size_t)rows * 32;
        if (bytes > zkh::kSlabBytes) return hipErrorInvalidValue;
        for (int c = 0; c < count; c++) cols[c] = nullptr;
        for (int c = 0; c < count; c++) {
            cols[c] = pool().alloc(bytes);
            if (!cols[c]) {
                for (int d = 0; d < c; d++) {
                    (void)pool().release(cols[d]);
                    cols[d] = nullptr;
                }
                return hipErrorOutOfMemory;
            }
        }
        unsigned hw = std::thread::hardware_concurrency();
        unsigned nt = hw ? (hw < 16 ? hw : 16) : 1;
        if (nt > (unsigned)count) nt = (unsigned)count;
        auto zero = [&](unsigned t) {
            for (int c = (int)t; c < count; c += (int)nt) memset(cols[c], 0, bytes);
        };
        std::vector<std::thread> th;
        unsigned started = 0;
        try {
            for (unsigned t = 1; t < nt; t++, started++) th.emplace_back(zero, t);
        } catch (...) {
        }
        zero(0);
        for (auto &x : th) x.join();
        for (unsigned t = started + 1; t < nt; t++) zero(t);  // shares whose thread could not be started
        return hipSuccess;
    });
}

hipError_t zk_release_advice_buffer(int count, void **cols) {
    if (count < 0 || (count && !cols)) return hipErrorInvalidValue;
    hipError_t err = hipSuccess;
    for (int c = 0; c < count; c++) {
        if (!cols[c]) continue;
        if (!pool().release(cols[c])) err = hipErrorInvalidValue;
        cols[c] = nullptr;
    }
    return err;
}

hipError_t zk_pinned_stats(size_t *slabs, size_t *chunks_in_use, size_t *slab_bytes, size_t *chunk_bytes) {
    pool().stats(slabs, chunks_in_use);
    if (slab_bytes) *slab_bytes = zkh::kSlabBytes;
    if (chunk_bytes) *chunk_bytes = zkh::kChunkBytes;
    return hipSuccess;
}

hipError_t zk_copyq_create(int max_inflight, void **out) {
    return zk::guard([&]() -> hipError_t {
        if (!out || max_inflight < 1) return hipErrorInvalidValue;
        auto *cq = new CopyQueue();
        cq->max = max_inflight;
        hipError_t err = hipStreamCreateWithFlags(&cq->cs, hipStreamNonBlocking);
        if (err != hipSuccess) {
            delete cq;
            return err;
        }
        *out = cq;
        return hipSuccess;
    });
}

hipError_t zk_copyq_push(void *q, void *host_dst, const void *dev_src, size_t bytes, hipStream_t producer,
                         int flags) {
    return zk::guard([&]() -> hipError_t {
        if (!q || !host_dst || !dev_src || bytes == 0) return hipErrorInvalidValue;
        if ((flags & ~(ZK_COPYQ_SNAPSHOT | ZK_COPYQ_RELEASE)) ||
            (flags & ZK_COPYQ_SNAPSHOT && flags & ZK_COPYQ_RELEASE))
            return hipErrorInvalidValue;
        CopyQueue &cq = *(CopyQueue *)q;
        std::lock_guard<std::mutex> g(cq.mu);
        hipError_t err = hipSuccess;
        if ((int)cq.q.size() >= cq.max) {
            auto t0 = std::chrono::steady_clock::now();
            Entry e = cq.q.front();
            cq.q.pop_front();
            err = finish(cq, e);
            double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            cq.full_waits++;
            if (ms > cq.max_wait_ms) cq.max_wait_ms = ms;
            if (err != hipSuccess) return err;
        }
        // the queue slot is taken before any work is enqueued: once a copy into
        // host_dst is in flight, nothing below can throw
        cq.q.emplace_back();
        Entry &e = cq.q.back();
        // once a copy is enqueued on cq.cs, a later failure waits for the
        // stream before the stage and the caller's buffer are given back
        bool enqueued = false;
    #define QCHECK(x)                                                                    \
        do {                                                                             \
            err = (x);                                                                   \
            if (err != hipSuccess) {                                                     \
                if (enqueued) (void)hipStreamSynchronize(cq.cs);                         \
                if (e.stage >= 0) cq.stages[e.stage].busy = false;                       \
                for (hipEvent_t ev : {e.ready, e.snap, e.done})                          \
                    if (ev) (void)hipEventDestroy(ev);                                   \
                cq.q.pop_back();                                                         \
                return err;                                                              \
            }                                                                            \
        } while (0)
        QCHECK(hipEventCreateWithFlags(&e.ready, hipEventDisableTiming));
        QCHECK(hipEventCreateWithFlags(&e.done, hipEventDisableTiming));
        QCHECK(hipEventRecord(e.ready, producer));
        QCHECK(hipStreamWaitEvent(cq.cs, e.ready, 0));
        const void *src = dev_src;
        if (flags & ZK_COPYQ_SNAPSHOT) {
            e.stage = take_stage(cq, bytes, &err);
            if (e.stage < 0) QCHECK(err);
            QCHECK(hipEventCreateWithFlags(&e.snap, hipEventDisableTiming));
            QCHECK(hipMemcpyAsync(cq.stages[e.stage].p, dev_src, bytes, hipMemcpyDeviceToDevice, cq.cs));
            enqueued = true;
            QCHECK(hipEventRecord(e.snap, cq.cs));
            QCHECK(hipStreamWaitEvent(producer, e.snap, 0));
            src = cq.stages[e.stage].p;
        }
        QCHECK(hipMemcpyAsync(host_dst, src, bytes, hipMemcpyDeviceToHost, cq.cs));
        enqueued = true;
        QCHECK(hipEventRecord(e.done, cq.cs));
    #undef QCHECK
        if (flags & ZK_COPYQ_RELEASE) e.release = (void *)dev_src;
        cq.pushed++;
        return hipSuccess;
    });
}

hipError_t zk_copyq_sync_all(void *q) {
    if (!q) return hipErrorInvalidValue;
    CopyQueue &cq = *(CopyQueue *)q;
    std::lock_guard<std::mutex> g(cq.mu);
    return drain(cq);
}

hipError_t zk_copyq_stats(void *q, unsigned long long *pushed, unsigned long long *full_waits, double *max_wait_ms) {
    if (!q) return hipErrorInvalidValue;
    CopyQueue &cq = *(CopyQueue *)q;
    std::lock_guard<std::mutex> g(cq.mu);
    if (pushed) *pushed = cq.pushed;
    if (full_waits) *full_waits = cq.full_waits;
    if (max_wait_ms) *max_wait_ms = cq.max_wait_ms;
    return hipSuccess;
}

hipError_t zk_copyq_destroy(void *q) {
    if (!q) return hipErrorInvalidValue;
    auto *cq = (CopyQueue *)q;
    hipError_t err;
    {
        std::lock_guard<std::mutex> g(cq->mu);
        err = drain(*cq);
        for (auto &s : cq->stages)
            if (s.p) (void)hipFree(s.p);
        (void)hipStreamDestroy(cq->cs);
    }
    delete cq;
    return err;
}

}  // extern "C"
