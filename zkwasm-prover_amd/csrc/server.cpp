// Proof server: continuous admission of proof requests of mixed circuit
// sizes onto one GPU.
//
// Each request names a circuit shape (one of the shapes the server was built
// with; their keys stay resident) and a witness seed.  `slots` worker threads,
// each with its own stream (default priority), take requests FIFO as slots
// and HBM free up: a request that does not fit yet does not block smaller
// ones behind it (backfill), until it has been passed over kMaxBypass times;
// then nothing starts ahead of it.  A second proof joins one in flight only
// when more than two requests wait per proof in flight (ZKS_OVERLAP_RATIO;
// with ZKS_EAGER: whenever a slot and a state are free), so at light load a
// proof runs alone, as on the one-at-a-time path, and under load proofs
// overlap.
// Per-proof device states are created when the server starts (one per shape,
// then up to `slots` per shape, smallest shapes first, within the HBM budget)
// and reused from per-shape pools: a request whose shape has no idle state
// waits for one, so no state is created or freed while proofs run.  Only if
// one state of every shape does not fit the budget are states created on
// demand and idle states of other shapes freed to make room.
//
// With ZKS_CROSS_BATCH the commitment MSMs of the proofs in flight go
// through one batcher: a group submitted while the batcher's stream is busy
// waits, and every group waiting when it frees up runs in ONE zk_msm_batch
// pass (items of several proofs and shapes, uneven sizes), after the
// batcher's stream has waited for each submitting proof's stream to reach
// its commit.  The reference proves one circuit at a time (src/lib.rs:205),
// so none of this has a counterpart there; each proof's arithmetic is
// zkp_prove, unchanged, so a result does not depend on what else runs.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "guard.h"
#include "zkinternal.h"
#include "zkwasm_prover_kernel.h"

namespace {

#define SCHECK(x)                             \
    do {                                      \
        hipError_t e_ = (x);                  \
        if (e_ != hipSuccess) return e_;      \
    } while (0)

constexpr int kMaxBypass = 8;        // later requests started past a waiting head
constexpr int kGroupItems = 16;      // MSM items of one prover-core commitment group

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct ShapeState {
    zkp_circuit circuit;
    void *pk = nullptr;
    size_t job_bytes = 0;
    int ncommit = 0;
    int states = 0;            // per-proof states of this shape, idle or in use
    std::vector<void *> idle;  // pooled per-proof states
};

struct Request {
    int shape = 0;
    uint64_t seed = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    int bypassed = 0;
    std::vector<uint8_t> commitments;
    zkp_proof_info info{};
    hipError_t err = hipSuccess;
    bool done = false;
};

// Cross-proof MSM batcher.  A caller's ticket waits until a leader (the first
// waiting caller while no batch runs) takes every pending ticket that fits
// the scratch and runs them as one zk_msm_batch pass on the batcher's stream.
struct Batcher {
    std::mutex mu;
    std::condition_variable cv;
    struct Ticket {
        const zk_msm_item *items;
        int count;
        uint8_t *out;
        hipEvent_t ready;
        bool done;
        hipError_t err;
    };
    std::deque<Ticket *> pending;
    bool running = false;
    hipStream_t st = nullptr;
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    uint8_t *out_dev = nullptr;
    int max_items = 0;
    unsigned long long launches = 0, merged = 0, items = 0;

    hipError_t init(size_t bytes, int max_it) {
        scratch_bytes = bytes;
        max_items = max_it;
        SCHECK(hipStreamCreate(&st));
        SCHECK(hipMalloc(&scratch, bytes));
        return hipMalloc((void **)&out_dev, (size_t)max_it * ZK_MSM_OUT_BYTES);
    }
    void release() {
        if (st) (void)hipStreamSynchronize(st);
        if (scratch) (void)hipFree(scratch);
        if (out_dev) (void)hipFree(out_dev);
        if (st) (void)hipStreamDestroy(st);
        scratch = nullptr;
        out_dev = nullptr;
        st = nullptr;
    }

    // one merged pass over the given tickets (no lock held)
    hipError_t run(const std::vector<Ticket *> &batch) {
        std::vector<zk_msm_item> all;
        for (Ticket *t : batch) {
            SCHECK(hipStreamWaitEvent(st, t->ready, 0));
            all.insert(all.end(), t->items, t->items + t->count);
        }
        SCHECK(zk_msm_batch(out_dev, all.data(), (int)all.size(), 0, scratch, scratch_bytes, st));
        std::vector<uint8_t> jac(all.size() * 96);
        for (size_t m = 0; m < all.size(); m++)
            SCHECK(hipMemcpyAsync(&jac[m * 96], out_dev + m * ZK_MSM_OUT_BYTES, 96, hipMemcpyDeviceToHost, st));
        SCHECK(hipStreamSynchronize(st));
        size_t m = 0;
        for (Ticket *t : batch)
            for (int i = 0; i < t->count; i++, m++) {
                const uint8_t *q = &jac[m * 96];
                bool id = true;
                for (int b = 64; b < 96; b++) id &= q[b] == 0;
                if (id)
                    memset(t->out + (size_t)i * 64, 0, 64);
                else
                    memcpy(t->out + (size_t)i * 64, q, 64);
            }
        return hipSuccess;
    }

    hipError_t commit(const zk_msm_item *it, int count, uint8_t *out, hipStream_t caller) {
        if (count <= 0 || count > max_items) return hipErrorInvalidValue;
        Ticket t{it, count, out, nullptr, false, hipSuccess};
        SCHECK(hipEventCreateWithFlags(&t.ready, hipEventDisableTiming));
        hipError_t e = hipEventRecord(t.ready, caller);
        if (e != hipSuccess) {
            (void)hipEventDestroy(t.ready);
            return e;
        }
        std::unique_lock<std::mutex> lk(mu);
        try {
            pending.push_back(&t);
        } catch (...) {
            lk.unlock();
            (void)hipEventDestroy(t.ready);
            return hipErrorOutOfMemory;
        }
        while (!t.done) {
            if (running) {
                cv.wait(lk);
                continue;
            }
            // lead: every pending ticket whose items still fit, in order
            running = true;
            std::vector<Ticket *> batch;
            size_t nitems = 0;
            hipError_t r = hipSuccess;
            try {
                std::vector<zk_msm_item> probe;
                for (auto p = pending.begin(); p != pending.end();) {
                    std::vector<zk_msm_item> next(probe);
                    next.insert(next.end(), (*p)->items, (*p)->items + (*p)->count);
                    const size_t need = zk_msm_batch_scratch_bytes(next.data(), (int)next.size(), 0);
                    if (!batch.empty() && ((int)next.size() > max_items || !need || need > scratch_bytes)) {
                        ++p;
                        continue;
                    }
                    batch.push_back(*p);
                    probe.swap(next);
                    p = pending.erase(p);
                }
                nitems = probe.size();
            } catch (...) {
                r = hipErrorOutOfMemory;
            }
            if (batch.empty()) {  // nothing taken: this ticket fails alone
                for (auto p = pending.begin(); p != pending.end(); ++p)
                    if (*p == &t) {
                        pending.erase(p);
                        break;
                    }
                t.err = r != hipSuccess ? r : hipErrorUnknown;
                t.done = true;
                running = false;
                cv.notify_all();
                break;
            }
            lk.unlock();
            if (r == hipSuccess) r = zk::guard([&] { return run(batch); });
            lk.lock();
            launches++;
            merged += batch.size() > 1;
            items += nitems;
            for (Ticket *b : batch) {
                b->err = r;
                b->done = true;
            }
            running = false;
            cv.notify_all();
        }
        e = t.err;
        lk.unlock();
        (void)hipEventDestroy(t.ready);
        return e;
    }
};

hipError_t batcher_commit(void *ctx, const void *items, int count, uint8_t *out, hipStream_t st) {
    return ((Batcher *)ctx)->commit((const zk_msm_item *)items, count, out, st);
}

struct Server {
    std::vector<ShapeState> shapes;
    // a deque: zks_submit appends while workers hold references to earlier
    // requests, and push_back on a deque never moves existing elements
    std::deque<Request> reqs;
    std::deque<int> queue;
    size_t budget = 0, used = 0, peak_used = 0;
    int flags = ZKP_BATCHED;
    bool eager = false;
    int overlap_ratio = 2;  // ZKS_OVERLAP_RATIO: waiting requests per running proof before another joins
    int running = 0, peak_running = 0;
    bool stop = false;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<std::thread> workers;
    std::vector<hipStream_t> streams;
    int device = 0;
    std::unique_ptr<Batcher> batcher;

    // under mu: the first queued request that can start now, with a job
    // reserved for it (popped from the pool, or bytes reserved for a new one)
    bool pick(int &rid, void *&job, bool &fresh, std::vector<void *> &to_free) {
        if (queue.empty()) return false;
        // overlap only on backlog: a proof joins those in flight when more
        // than overlap_ratio requests wait per proof running
        if (!eager && running > 0 && (int)queue.size() <= running * overlap_ratio) return false;
        const bool head_only = reqs[queue.front()].bypassed >= kMaxBypass;
        for (auto it = queue.begin(); it != queue.end(); ++it) {
            if (head_only && it != queue.begin()) break;
            ShapeState &sh = shapes[reqs[*it].shape];
            if (!sh.idle.empty()) {
                job = sh.idle.back();
                sh.idle.pop_back();
                fresh = false;
            } else {
                // a shape that already has a state waits for it rather than
                // evicting the pooled states of other shapes (creating a
                // state of tens of GB next to running proofs takes seconds)
                if (sh.states > 0 && used + sh.job_bytes > budget) continue;
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
                            o.states--;
                            used -= o.job_bytes;
                        }
                    }
                }
                used += sh.job_bytes;
                if (used > peak_used) peak_used = used;
                sh.states++;
                job = nullptr;
                fresh = true;
            }
            if (it != queue.begin()) reqs[queue.front()].bypassed++;
            rid = *it;
            queue.erase(it);
            return true;
        }
        return false;
    }

    // The per-proof states, created before any request runs: one of every
    // shape (largest first), then more, smallest shapes first, up to `slots`
    // of a shape while the budget holds them.  Serving then allocates nothing
    // while the budget is full; if one state of every shape does not fit,
    // states are created and evicted on demand instead.
    hipError_t prefill(int slots) {
        std::vector<int> order(shapes.size());
        for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
        std::sort(order.begin(), order.end(),
                  [&](int a, int b) { return shapes[a].job_bytes > shapes[b].job_bytes; });
        size_t one_each = 0;
        for (auto &sh : shapes) one_each += sh.job_bytes;
        if (one_each > budget) return hipSuccess;
        auto add = [&](ShapeState &sh) -> hipError_t {
            void *job = nullptr;
            SCHECK(zkp_job_create(sh.pk, &job));
            sh.idle.push_back(job);  // capacity slots + 1 reserved
            sh.states++;
            used += sh.job_bytes;
            return hipSuccess;
        };
        for (int i : order) SCHECK(add(shapes[i]));
        for (auto it = order.rbegin(); it != order.rend(); ++it) {
            ShapeState &sh = shapes[*it];
            while (sh.states < slots && used + sh.job_bytes <= budget) SCHECK(add(sh));
        }
        if (used > peak_used) peak_used = used;
        return hipSuccess;
    }

    // under mu: a finished or failed proof gives back its slot and its job
    void finish(Request &rq, ShapeState &sh, void *job, hipError_t e) {
        running--;
        rq.err = e;
        rq.t_done = now_ms();
        rq.done = true;
        if (job && e == hipSuccess && sh.idle.size() < sh.idle.capacity()) {
            sh.idle.push_back(job);  // capacity reserved at creation: cannot throw
        } else {
            if (job) (void)zkp_job_destroy(job);
            used -= sh.job_bytes;
            sh.states--;
        }
        cv_done.notify_all();
        cv_work.notify_all();
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
                bool got = false;
                try {
                    to_free.reserve(64);
                    cv_work.wait(lk, [&] { return stop || (got = pick(rid, job, fresh, to_free)); });
                } catch (...) {
                    got = false;  // a host allocation failed: nothing was taken from the queue
                }
                if (!got) {
                    if (stop) return;
                    continue;
                }
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
            if (e == hipSuccess && batcher) e = zk::job_set_commit_fn(job, batcher_commit, batcher.get());
            if (e == hipSuccess) e = zkp_prove(job, rq.seed, flags, rq.commitments.data(), nullptr, &rq.info, st);
            if (job && batcher) (void)zk::job_set_commit_fn(job, nullptr, nullptr);
            std::lock_guard<std::mutex> lk(mu);
            finish(rq, sh, job, e);
        }
    }
};

// scratch of one proof's commitment group: plain points or the key's
// fixed-base tables, whichever needs more
size_t group_scratch(const zkp_circuit &c) {
    std::vector<zk_msm_item> it(kGroupItems, zk_msm_item{nullptr, nullptr, 1 << c.k, 0});
    const size_t plain = zk_msm_batch_scratch_bytes(it.data(), kGroupItems, 0);
    for (auto &x : it) x.table_c = zk_msm_table_window(1 << c.k);
    const size_t tab = zk_msm_batch_scratch_bytes(it.data(), kGroupItems, 0);
    return tab > plain ? tab : plain;
}

}  // namespace

extern "C" {

// Builds the keys of every shape (resident) and starts `slots` workers.
// hbm_budget: bytes the per-proof states may take together (0: what the
// device reports free after the keys and the batcher, less 4 GiB).
hipError_t zks_create(const zkp_circuit *shapes, int nshapes, int slots, size_t hbm_budget, int flags,
                      unsigned long long key_seed, void **srv_out) {
    return zk::guard([&]() -> hipError_t {
        if (!shapes || nshapes <= 0 || slots <= 0 || slots > 32 || !srv_out) return hipErrorInvalidValue;
        std::unique_ptr<Server> s(new Server());
        s->flags = flags & (ZKP_BATCHED | ZKP_SHPLONK | ZKP_NO_OPEN);
        s->eager = (flags & ZKS_EAGER) != 0;
        if (const char *r = getenv("ZKS_OVERLAP_RATIO")) {
            const int v = atoi(r);
            if (v >= 1 && v <= 64) s->overlap_ratio = v;
        }
        hipError_t e = hipGetDevice(&s->device);
        s->shapes.resize(nshapes);
        size_t max_group = 0;
        for (int i = 0; i < nshapes && e == hipSuccess; i++) {
            ShapeState &sh = s->shapes[i];
            sh.circuit = shapes[i];
            sh.job_bytes = zkp_job_bytes(&shapes[i]);
            sh.ncommit = zkp_commitment_count(&shapes[i]);
            if (!sh.job_bytes) e = hipErrorInvalidValue;
            if (e == hipSuccess) {
                sh.idle.reserve((size_t)slots + 1);
                const size_t g = group_scratch(shapes[i]);
                if (g > max_group) max_group = g;
                e = zkp_pk_create(&shapes[i], key_seed + (unsigned long long)i, nullptr, &sh.pk);
            }
        }
        if (e == hipSuccess && (flags & ZKS_CROSS_BATCH) && (s->flags & ZKP_BATCHED)) {
            // room for two commitment groups of the largest shape in one pass
            s->batcher.reset(new Batcher());
            e = s->batcher->init(2 * max_group + 256, 2 * kGroupItems);
        }
        if (e == hipSuccess) {
            size_t free_b = 0, total_b = 0;
            e = hipMemGetInfo(&free_b, &total_b);
            size_t avail = free_b > ((size_t)4 << 30) ? free_b - ((size_t)4 << 30) : 0;
            s->budget = hbm_budget && hbm_budget < avail ? hbm_budget : avail;
            for (auto &sh : s->shapes)
                if (sh.job_bytes > s->budget) e = hipErrorOutOfMemory;  // a shape that could never run
        }
        if (e == hipSuccess) e = s->prefill(slots);
        if (e != hipSuccess) {
            for (auto &sh : s->shapes) {
                for (void *j : sh.idle) (void)zkp_job_destroy(j);
                if (sh.pk) (void)zkp_pk_destroy(sh.pk);
            }
            if (s->batcher) s->batcher->release();
            return e;
        }
        s->streams.assign(slots, nullptr);
        for (int w = 0; w < slots && e == hipSuccess; w++) e = hipStreamCreate(&s->streams[w]);
        Server *raw = s.release();  // zks_destroy owns it from here
        if (e == hipSuccess) {
            try {
                for (int w = 0; w < slots; w++) raw->workers.emplace_back([raw, w] { raw->worker(w); });
            } catch (...) {
                e = hipErrorUnknown;  // a worker thread could not be started
            }
        }
        if (e != hipSuccess) {
            (void)zks_destroy(raw);  // stops and joins the workers started, frees streams and keys
            return e;
        }
        *srv_out = raw;
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

// Cross-proof batcher counters: merged passes run, passes that held groups
// of more than one proof, MSM items committed through it (zeros without
// ZKS_CROSS_BATCH).
hipError_t zks_batch_stats(void *srv, unsigned long long *launches, unsigned long long *merged,
                           unsigned long long *items) {
    Server *s = (Server *)srv;
    if (!s) return hipErrorInvalidValue;
    unsigned long long l = 0, m = 0, i = 0;
    if (s->batcher) {
        std::lock_guard<std::mutex> lk(s->batcher->mu);
        l = s->batcher->launches;
        m = s->batcher->merged;
        i = s->batcher->items;
    }
    if (launches) *launches = l;
    if (merged) *merged = m;
    if (items) *items = i;
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
        if (s->batcher) s->batcher->release();
        delete s;
        return hipSuccess;
    });
}

}  // extern "C"
