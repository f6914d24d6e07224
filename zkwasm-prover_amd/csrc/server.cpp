// Proof server: continuous admission of proof requests of mixed circuit
// sizes onto one GPU.
//
// Each request names a circuit shape (one of the shapes the server was built
// with; their keys stay resident) and a witness seed.  `slots` worker threads,
// each with its own stream (default priority), take requests as soon as a slot
// is free and the request's per-proof device state fits the HBM budget: FIFO
// order, but a request that does not fit yet does not block smaller ones
// behind it (backfill).  Per-proof device states are kept in per-shape pools
// and reused; idle states of other shapes are freed when a request needs the
// room.  Requests in flight overlap on the device, since their kernels go to
// different streams.  The reference proves one circuit at a time
// (src/lib.rs:205), so this has no counterpart there; each proof's arithmetic
// is zkp_prove, unchanged, so a result does not depend on what else runs.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "guard.h"
#include "zkwasm_prover_kernel.h"

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct ShapeState {
    zkp_circuit circuit;
    void *pk = nullptr;
    size_t job_bytes = 0;
    int ncommit = 0;
    std::vector<void *> idle;  // pooled per-proof states
};

struct Request {
    int shape = 0;
    uint64_t seed = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    std::vector<uint8_t> commitments;
    zkp_proof_info info{};
    hipError_t err = hipSuccess;
    bool done = false;
};

struct Server {
    std::vector<ShapeState> shapes;
    // a deque: zks_submit appends while workers hold references to earlier
    // requests, and push_back on a deque never moves existing elements
    std::deque<Request> reqs;
    std::deque<int> queue;
    size_t budget = 0, used = 0, peak_used = 0;
    int flags = ZKP_BATCHED;
    int running = 0, peak_running = 0;
    bool stop = false;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<std::thread> workers;
    std::vector<hipStream_t> streams;
    int device = 0;

    // under mu: the first queued request that can start now, with a job
    // reserved for it (popped from the pool, or bytes reserved for a new one)
    bool pick(int &rid, void *&job, bool &fresh, std::vector<void *> &to_free) {
        for (auto it = queue.begin(); it != queue.end(); ++it) {
            ShapeState &sh = shapes[reqs[*it].shape];
            if (!sh.idle.empty()) {
                job = sh.idle.back();
                sh.idle.pop_back();
                fresh = false;
            } else {
                if (used + sh.job_bytes > budget) {
                    // free idle states of other shapes until it fits
                    size_t idle_other = 0;
                    for (auto &o : shapes)
                        if (&o != &sh) idle_other += o.idle.size() * o.job_bytes;
                    if (used - idle_other + sh.job_bytes > budget) continue;
                    for (auto &o : shapes) {
                        while (&o != &sh && !o.idle.empty() && used + sh.job_bytes > budget) {
                            to_free.push_back(o.idle.back());
                            o.idle.pop_back();
                            used -= o.job_bytes;
                        }
                    }
                }
                used += sh.job_bytes;
                if (used > peak_used) peak_used = used;
                job = nullptr;
                fresh = true;
            }
            rid = *it;
            queue.erase(it);
            return true;
        }
        return false;
    }

    void worker(int w) {
        (void)hipSetDevice(device);
        hipStream_t st = streams[w];
        for (;;) {
            int rid = -1;
            void *job = nullptr;
            bool fresh = false;
            std::vector<void *> to_free;
            Request *req = nullptr;  // element addresses of the deque are stable
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stop || pick(rid, job, fresh, to_free); });
                if (rid < 0) return;  // stop with nothing admissible
                running++;
                if (running > peak_running) peak_running = running;
                req = &reqs[rid];
                req->t_start = now_ms();
            }
            for (void *j : to_free) (void)zkp_job_destroy(j);
            Request &rq = *req;
            ShapeState &sh = shapes[rq.shape];
            hipError_t e = hipSuccess;
            if (fresh) e = zkp_job_create(sh.pk, &job);
            if (e == hipSuccess)
                e = zkp_prove(job, rq.seed, flags, rq.commitments.data(), nullptr, &rq.info, st);
            {
                std::lock_guard<std::mutex> lk(mu);
                running--;
                rq.err = e;
                rq.t_done = now_ms();
                rq.done = true;
                if (job && e == hipSuccess) {
                    sh.idle.push_back(job);
                } else {
                    if (job) (void)zkp_job_destroy(job);
                    used -= sh.job_bytes;
                }
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

}  // namespace

extern "C" {

// Builds the keys of every shape (resident) and starts `slots` workers.
// hbm_budget: bytes the per-proof states may take together (0: what the
// device reports free after the keys, less 4 GiB).
hipError_t zks_create(const zkp_circuit *shapes, int nshapes, int slots, size_t hbm_budget, int flags,
                      unsigned long long key_seed, void **srv_out) {
    return zk::guard([&]() -> hipError_t {
        if (!shapes || nshapes <= 0 || slots <= 0 || slots > 32 || !srv_out) return hipErrorInvalidValue;
        Server *s = new Server();
        s->flags = flags & (ZKP_BATCHED | ZKP_SHPLONK | ZKP_NO_OPEN);
        hipError_t e = hipGetDevice(&s->device);
        s->shapes.resize(nshapes);
        for (int i = 0; i < nshapes && e == hipSuccess; i++) {
            ShapeState &sh = s->shapes[i];
            sh.circuit = shapes[i];
            sh.job_bytes = zkp_job_bytes(&shapes[i]);
            sh.ncommit = zkp_commitment_count(&shapes[i]);
            if (!sh.job_bytes) e = hipErrorInvalidValue;
            if (e == hipSuccess) e = zkp_pk_create(&shapes[i], key_seed + (unsigned long long)i, nullptr, &sh.pk);
        }
        if (e == hipSuccess) {
            size_t free_b = 0, total_b = 0;
            e = hipMemGetInfo(&free_b, &total_b);
            size_t avail = free_b > ((size_t)4 << 30) ? free_b - ((size_t)4 << 30) : 0;
            s->budget = hbm_budget && hbm_budget < avail ? hbm_budget : avail;
            for (auto &sh : s->shapes)
                if (sh.job_bytes > s->budget) e = hipErrorOutOfMemory;  // a shape that could never run
        }
        if (e != hipSuccess) {
            for (auto &sh : s->shapes)
                if (sh.pk) (void)zkp_pk_destroy(sh.pk);
            delete s;
            return e;
        }
        s->streams.assign(slots, nullptr);
        for (int w = 0; w < slots && e == hipSuccess; w++) e = hipStreamCreate(&s->streams[w]);
        if (e == hipSuccess) {
            try {
                for (int w = 0; w < slots; w++) s->workers.emplace_back([s, w] { s->worker(w); });
            } catch (...) {
                e = hipErrorUnknown;  // a worker thread could not be started
            }
        }
        if (e != hipSuccess) {
            (void)zks_destroy(s);  // stops and joins the workers started, frees streams and keys
            return e;
        }
        *srv_out = s;
        return hipSuccess;
    });
}

// Queues one proof; returns its request id (>= 0) or -1.
int zks_submit(void *srv, int shape, unsigned long long seed) {
    return zk::guard_or<int>(-1, [&]() -> int {
        Server *s = (Server *)srv;
        if (!s || shape < 0 || shape >= (int)s->shapes.size()) return -1;
        // built outside the list: an allocation that fails leaves no request
        // behind that would never complete (zks_destroy waits for all of them)
        Request r;
        r.shape = shape;
        r.seed = seed;
        r.commitments.assign((size_t)s->shapes[shape].ncommit * 64, 0);
        int id;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            id = (int)s->reqs.size();
            r.t_submit = now_ms();
            s->reqs.push_back(std::move(r));  // if this throws, nothing was added
            try {
                s->queue.push_back(id);
            } catch (...) {
                s->reqs.pop_back();
                throw;
            }
        }
        s->cv_work.notify_all();
        return id;
    });
}

// Waits for request id; copies its commitments (ncommit * 64 B, may be null)
// and its result record.
hipError_t zks_wait(void *srv, int id, void *commitments, zks_result *res) {
    return zk::guard([&]() -> hipError_t {
        Server *s = (Server *)srv;
        if (!s) return hipErrorInvalidValue;
        std::unique_lock<std::mutex> lk(s->mu);
        if (id < 0 || id >= (int)s->reqs.size()) return hipErrorInvalidValue;
        s->cv_done.wait(lk, [&] { return s->reqs[id].done; });
        Request &r = s->reqs[id];
        if (commitments) memcpy(commitments, r.commitments.data(), r.commitments.size());
        if (res) {
            memset(res, 0, sizeof(*res));
            res->shape = r.shape;
            res->err = (int)r.err;
            res->z_ok = r.info.z_ok;
            res->quotient_ok = r.info.quotient_ok;
            res->ncommit = r.info.ncommit;
            res->queue_ms = r.t_start - r.t_submit;
            res->run_ms = r.t_done - r.t_start;
            res->latency_ms = r.t_done - r.t_submit;
            res->prove_ms = r.info.total_ms;
        }
        return r.err;
    });
}

// Server-wide counters: peak proofs in flight, peak and budget bytes.
hipError_t zks_stats(void *srv, int *peak_running, size_t *peak_bytes, size_t *budget) {
    Server *s = (Server *)srv;
    if (!s) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lk(s->mu);
    if (peak_running) *peak_running = s->peak_running;
    if (peak_bytes) *peak_bytes = s->peak_used;
    if (budget) *budget = s->budget;
    return hipSuccess;
}

// Finishes every queued request, stops the workers, frees all device state.
hipError_t zks_destroy(void *srv) {
    return zk::guard([&]() -> hipError_t {
        Server *s = (Server *)srv;
        if (!s) return hipSuccess;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_done.wait(lk, [&] {
                for (auto &r : s->reqs)
                    if (!r.done) return false;
                return true;
            });
            s->stop = true;
        }
        s->cv_work.notify_all();
        for (auto &t : s->workers) t.join();
        for (auto st : s->streams)
            if (st) (void)hipStreamDestroy(st);
        for (auto &sh : s->shapes) {
            for (void *j : sh.idle) (void)zkp_job_destroy(j);
            if (sh.pk) (void)zkp_pk_destroy(sh.pk);
        }
        delete s;
        return hipSuccess;
    });
}

}  // extern "C"
