// Reconstruction server: continuous batching of McSART jobs on one GPU.
//
// No counterpart in the reference, whose driver reconstructs one scan per MATLAB call
// (src/entry_AUMISART.cu:193-207).  Requests queue in arrival order.  A request is admitted as
// soon as a slot is free and its device footprint (mcsart_job_bytes) fits in the HBM budget;
// admission runs under the server's lock in whichever thread frees a slot or submits a request.
//
// Each slot is a stream (its own context) driven by its own host thread, so one slot's uploads,
// result copies or waits never hold up another slot.  An admitted request becomes a
// device-resident job in a block of device memory from the server's pool: inputs uploaded once,
// then the schedule's sweeps enqueued kernel by kernel (each runs once, so a captured graph would
// never be replayed) with at most two sweeps in flight (the thread waits for sweep k-1 before
// enqueueing sweep k+1), so a cancellation takes effect between sweeps while the stream never
// runs dry.  When the last sweep completes the results go into the caller's buffers and the block
// returns to the pool.
//
// Block pool: a finished job's block is kept for a later request of the same or somewhat smaller
// footprint (up to twice the request's bytes) instead of being freed, because hipFree waits for
// the whole device: freeing on the completion path would hold every slot's thread until all
// in-flight jobs finished, turning continuous batching into lockstep gangs.  Idle blocks are
// freed only when the budget needs their memory, and at destroy.
//
// A job's kernels, launch shapes and their order do not depend on what else runs, so a request's
// results are the same alone and in any mix.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "internal.hpp"

using namespace mcsart;
using Clock = std::chrono::steady_clock;

namespace {

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Block {
    void *ptr = nullptr;
    size_t bytes = 0;
};

struct Request {
    long long id = 0;
    mcsart_request in{};
    std::vector<float> angles, volumes, flows;   // small per-view arrays are copied at submit
    std::vector<int> op_iter;
    size_t bytes = 0;
    int state = MCSART_REQ_QUEUED;
    int status = MCSART_OK;
    std::string error;   // the failure's message, handed to the waiting thread
    int slot = -1;
    int bypassed = 0;   // times a later request was admitted ahead of this one
    std::atomic<bool> cancel{false};
    mcsart_tuning tuning{0, 0, 0, 0, 0};
    // the device block it runs in: reused from the pool, or allocated by the slot's thread
    Block block;
    bool allocate = false;
    std::vector<Block> to_free;   // idle blocks the admission decided to give back first
    Clock::time_point t_submit, t_admit, t_done;
};

struct Slot {
    mcsart_ctx *ctx = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};   // completion of the last two enqueued sweeps
    Request *req = nullptr;                  // set under the lock by admission
    std::condition_variable cv;
    std::thread th;
};

}  // namespace

struct mcsart_server {
    int device = 0;
    mcsart_server_opts opts{};
    size_t budget = 0;
    size_t used = 0;             // device bytes the server holds: running jobs' and idle blocks
    std::vector<Block> idle;     // the pool
    std::vector<Slot *> slots;
    std::deque<Request *> queue;
    std::map<long long, Request *> reqs;
    long long next_id = 1;
    std::mutex mu;
    std::condition_variable cv_done;   // a request ended (done, failed or cancelled)
    bool stop = false;
    std::string worker_error;          // a slot thread could not start
    Clock::time_point t0 = Clock::now();   // origin of the requests' admit/done times
};

namespace {

size_t idle_bytes(const mcsart_server *sv) {
    size_t b = 0;
    for (const Block &x : sv->idle) b += x.bytes;
    return b;
}

// The idle block that best fits `bytes` (smallest of at least `bytes` and at most twice that),
// or -1.
int best_idle(const mcsart_server *sv, size_t bytes) {
    int best = -1;
    for (int k = 0; k < (int)sv->idle.size(); ++k) {
        const size_t b = sv->idle[k].bytes;
        if (b >= bytes && b <= 2 * bytes && (best < 0 || b < sv->idle[best].bytes)) best = k;
    }
    return best;
}

// Whether a request of `bytes` can be admitted now: a pooled block fits it, or a new block fits
// the budget once idle blocks are given back.
bool fits(const mcsart_server *sv, size_t bytes) {
    return best_idle(sv, bytes) >= 0 || sv->used - idle_bytes(sv) + bytes <= sv->budget;
}

// Claim device memory for r (under the lock): a pooled block, or the right to allocate a new
// one after giving back the largest idle blocks the budget needs.
void claim(mcsart_server *sv, Request *r) {
    const int k = best_idle(sv, r->bytes);
    if (k >= 0) {
        r->block = sv->idle[k];
        r->allocate = false;
        sv->idle.erase(sv->idle.begin() + k);
        return;
    }
    std::sort(sv->idle.begin(), sv->idle.end(),
              [](const Block &a, const Block &b) { return a.bytes > b.bytes; });
    while (sv->used + r->bytes > sv->budget && !sv->idle.empty()) {
        r->to_free.push_back(sv->idle.front());
        sv->used -= sv->idle.front().bytes;
        sv->idle.erase(sv->idle.begin());
    }
    r->block = Block{nullptr, r->bytes};
    r->allocate = true;
    sv->used += r->bytes;
}

// First request in arrival order that fits.  A request that fits nowhere yet is bypassed at most
// 2*slots times; after that nothing is admitted ahead of it (no starvation).
Request *pick(mcsart_server *sv) {
    for (size_t k = 0; k < sv->queue.size(); ++k) {
        Request *r = sv->queue[k];
        if (fits(sv, r->bytes)) {
            for (size_t m = 0; m < k; ++m) sv->queue[m]->bypassed++;
            sv->queue.erase(sv->queue.begin() + (long)k);
            return r;
        }
        if (r->bypassed >= 2 * (int)sv->slots.size()) return nullptr;
    }
    return nullptr;
}

// Give every free slot a request that fits (called with the lock held).
void admit_locked(mcsart_server *sv) {
    for (Slot *s : sv->slots) {
        if (s->req) continue;
        Request *r = pick(sv);
        if (!r) return;
        claim(sv, r);
        r->state = MCSART_REQ_RUNNING;
        r->slot = (int)(std::find(sv->slots.begin(), sv->slots.end(), s) - sv->slots.begin());
        r->t_admit = Clock::now();
        s->req = r;
        s->cv.notify_one();
    }
}

// The request's whole life on slot s, outside the lock: memory, job, uploads, sweeps (at most two
// in flight, stopping early on cancellation), results.  Returns the status.
int run_request(mcsart_server *sv, Slot *s, Request *r) {
    for (const Block &b : r->to_free) (void)hipFree(b.ptr);
    r->to_free.clear();
    if (r->allocate) {
        if (hipMalloc(&r->block.ptr, r->block.bytes) != hipSuccess) {
            (void)hipGetLastError();
            r->block.ptr = nullptr;
            return fail(MCSART_ENOMEM, "allocation of " + std::to_string(r->block.bytes) +
                                           " bytes for a job failed");
        }
    }
    const mcsart_request &q = r->in;
    mcsart_job *j = nullptr;
    int rc = job_create_in(s->ctx, &q.geom, q.n_views, r->angles.data(), r->volumes.data(),
                           r->flows.data(), &q.opts, q.n_iter, r->op_iter.data(), r->block.ptr,
                           r->block.bytes, &j);
    if (rc) return rc;
    auto body = [&]() -> int {
        int rc2;
        // Each sweep of a served request runs once, so a captured graph would never be replayed:
        // the sweeps are launched kernel by kernel.  (Stream capture is also kept out of the slot
        // threads, which run concurrently.)
        if ((rc2 = mcsart_job_set_graphs(j, 0))) return rc2;
        if ((rc2 = mcsart_job_set_projections(j, q.h_proj))) return rc2;
        if ((rc2 = mcsart_job_set_image(j, q.h_img))) return rc2;
        if ((rc2 = mcsart_job_set_fields(j, q.h_fields))) return rc2;   // NULL components: zero
        if (sv->opts.autotune) {
            if ((rc2 = mcsart_job_autotune(j, &r->tuning))) return rc2;
        } else {
            const mcsart_tuning t{sv->opts.proj_split, sv->opts.back_zslab, sv->opts.back_stage,
                                  sv->opts.proj_rows, sv->opts.back_zslab_motion};
            if ((rc2 = mcsart_job_set_tuning(j, &t))) return rc2;
            r->tuning = t;
        }
        for (int it = 0; it < q.n_iter; ++it) {
            if (r->cancel.load()) break;
            if ((rc2 = mcsart_job_run_sweep(j, it))) return rc2;
            MCSART_HIP_TRY(hipEventRecord(s->ev[it & 1], s->ctx->stream));
            if (it >= 1) MCSART_HIP_TRY(hipEventSynchronize(s->ev[(it - 1) & 1]));
        }
        MCSART_HIP_TRY(hipStreamSynchronize(s->ctx->stream));
        if (r->cancel.load()) return MCSART_ECANCELED;
        if (q.out_img && (rc2 = mcsart_job_get_image(j, q.out_img))) return rc2;
        if (q.out_norm && (rc2 = mcsart_job_get_norms(j, q.out_norm))) return rc2;
        bool any = false;
        for (int c = 0; c < 6; ++c) any |= q.out_fields[c] != nullptr;
        if (any && (rc2 = mcsart_job_get_fields(j, q.out_fields))) return rc2;
        return MCSART_OK;
    };
    rc = body();
    mcsart_job_destroy(j);   // waits for the slot's stream; the block stays allocated
    return rc;
}

void slot_main(mcsart_server *sv, Slot *s) {
    if (hipSetDevice(sv->device) != hipSuccess) {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->worker_error = "hipSetDevice failed in a slot thread";
        sv->cv_done.notify_all();
        return;
    }
    std::unique_lock<std::mutex> lk(sv->mu);
    while (true) {
        s->cv.wait(lk, [&] { return s->req != nullptr || (sv->stop && sv->queue.empty()); });
        if (!s->req) return;   // stopping and nothing left to run
        Request *r = s->req;
        lk.unlock();
        const int rc = run_request(sv, s, r);
        const std::string msg = rc && rc != MCSART_ECANCELED ? mcsart_last_error() : "";
        lk.lock();
        if (r->block.ptr) {
            sv->idle.push_back(r->block);   // back to the pool
        } else if (r->allocate) {
            sv->used -= r->block.bytes;     // the allocation failed: nothing is held
        }
        r->block = Block{};
        r->t_done = Clock::now();
        r->status = rc;
        r->error = rc == MCSART_ECANCELED ? "request cancelled" : msg;
        r->state = rc == MCSART_OK          ? MCSART_REQ_DONE
                   : rc == MCSART_ECANCELED ? MCSART_REQ_CANCELLED
                                            : MCSART_REQ_FAILED;
        s->req = nullptr;
        sv->cv_done.notify_all();
        admit_locked(sv);
        if (sv->stop && sv->queue.empty()) {
            for (Slot *o : sv->slots) o->cv.notify_one();   // idle slots may exit now
        }
    }
}

bool ended(const Request *r) {
    return r->state == MCSART_REQ_DONE || r->state == MCSART_REQ_FAILED ||
           r->state == MCSART_REQ_CANCELLED;
}

int check_request(const mcsart_request *q) {
    if (!q) return fail(MCSART_EINVAL, "null request");
    if (!q->angles || !q->volumes || !q->flows || !q->h_proj)
        return fail(MCSART_EINVAL, "request without angles, volumes, flows or projections");
    if (q->n_views <= 0) return fail(MCSART_EINVAL, "n_views must be positive");
    if (q->n_iter < 0 || (q->n_iter > 0 && !q->op_iter))
        return fail(MCSART_EINVAL, "bad op_iter");
    const mcsart_geom &g = q->geom;
    const Geom G{g.nx, g.ny, g.nz, g.na, g.nb, g.da, g.db, g.ai, g.bi, g.SO, g.SD};
    int rc = validate_geom(G);
    if (rc) return rc;
    return validate_volume(g.nx, g.ny, g.nz, true);
}

void destroy_slots(mcsart_server *sv) {
    for (Slot *s : sv->slots) {
        if (s->th.joinable()) s->th.join();
        for (hipEvent_t e : s->ev)
            if (e) (void)hipEventDestroy(e);
        if (s->ctx) mcsart_ctx_destroy(s->ctx);
        delete s;
    }
    sv->slots.clear();
}

}  // namespace

extern "C" int mcsart_server_create(int device, const mcsart_server_opts *o,
                                    mcsart_server **out) {
    if (!out) return fail(MCSART_EINVAL, "null server pointer");
    *out = nullptr;
    if (!o) return fail(MCSART_EINVAL, "null options");
    if (o->slots < 1 || o->slots > 16) return fail(MCSART_EINVAL, "slots must be in 1..16");
    const mcsart_tuning t{o->proj_split, o->back_zslab, o->back_stage, o->proj_rows,
                          o->back_zslab_motion};
    if (t.proj_split < 0 || t.proj_split > 16 || (t.proj_split & (t.proj_split - 1)) ||
        (t.proj_rows != 0 && t.proj_rows != 1 && t.proj_rows != 2 && t.proj_rows != 4) ||
        t.back_zslab < 0 || t.back_zslab > 64 || t.back_stage < 0 || t.back_stage > 2 ||
        t.back_zslab_motion < 0 || t.back_zslab_motion > 64)
        return fail(MCSART_EINVAL, "bad fixed launch shapes");
    MCSART_HIP_TRY(hipSetDevice(device));
    auto *sv = new mcsart_server();
    sv->device = device;
    sv->opts = *o;
    size_t fb = 0, tb = 0;
    if (hipMemGetInfo(&fb, &tb) != hipSuccess) {
        delete sv;
        return fail(MCSART_EHIP, "hipMemGetInfo failed");
    }
    sv->budget = o->hbm_budget ? std::min(o->hbm_budget, fb) : (size_t)(0.9 * (double)fb);
    for (int k = 0; k < o->slots; ++k) {
        auto *s = new Slot();
        sv->slots.push_back(s);
        int rc = mcsart_ctx_create(&s->ctx, device);
        for (int e = 0; e < 2 && !rc; ++e)
            if (hipEventCreateWithFlags(&s->ev[e], hipEventDisableTiming | hipEventBlockingSync) !=
                hipSuccess)
                rc = fail(MCSART_EHIP, "hipEventCreate failed");
        if (rc) {
            destroy_slots(sv);
            delete sv;
            return rc;
        }
    }
    for (Slot *s : sv->slots) s->th = std::thread(slot_main, sv, s);
    *out = sv;
    return MCSART_OK;
}

extern "C" int mcsart_server_submit(mcsart_server *sv, const mcsart_request *q, long long *id) {
    if (!sv) return fail(MCSART_EINVAL, "null server");
    int rc = check_request(q);
    if (rc) return rc;
    const size_t bytes = mcsart_job_bytes(&q->geom, q->n_views, q->n_iter, q->opts.keep_residual);
    if (bytes == 0 || bytes > sv->budget)
        return fail(MCSART_ENOMEM, "request needs " + std::to_string(bytes) +
                                       " bytes of device memory; the budget is " +
                                       std::to_string(sv->budget));
    auto *r = new Request();
    r->in = *q;
    r->angles.assign(q->angles, q->angles + q->n_views);
    r->volumes.assign(q->volumes, q->volumes + q->n_views);
    r->flows.assign(q->flows, q->flows + q->n_views);
    r->op_iter.assign(q->op_iter, q->op_iter + q->n_iter);
    r->bytes = bytes;
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stop) {
        delete r;
        return fail(MCSART_EBUSY, "server is shutting down");
    }
    r->id = sv->next_id++;
    sv->reqs[r->id] = r;
    sv->queue.push_back(r);
    if (id) *id = r->id;
    admit_locked(sv);
    return MCSART_OK;
}

extern "C" int mcsart_server_cancel(mcsart_server *sv, long long id) {
    if (!sv) return fail(MCSART_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return fail(MCSART_EINVAL, "unknown request id");
    Request *r = it->second;
    if (ended(r)) return fail(MCSART_EINVAL, "request " + std::to_string(id) + " has ended");
    if (r->state == MCSART_REQ_QUEUED) {   // never admitted: it leaves the queue now
        sv->queue.erase(std::find(sv->queue.begin(), sv->queue.end(), r));
        r->state = MCSART_REQ_CANCELLED;
        r->status = MCSART_ECANCELED;
        r->error = "request cancelled";
        r->t_done = Clock::now();
        sv->cv_done.notify_all();
        admit_locked(sv);   // a request that waited behind it may fit now
        return MCSART_OK;
    }
    r->cancel.store(true);   // running: its slot stops after the sweeps already enqueued
    return MCSART_OK;
}

extern "C" int mcsart_server_wait(mcsart_server *sv, long long id, double timeout_s) {
    if (!sv) return fail(MCSART_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return fail(MCSART_EINVAL, "unknown request id");
    Request *r = it->second;
    auto done = [&] { return ended(r) || !sv->worker_error.empty(); };
    if (timeout_s < 0) {
        sv->cv_done.wait(lk, done);
    } else if (!sv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) {
        return fail(MCSART_EBUSY, "request " + std::to_string(id) + " still running");
    }
    if (r->state == MCSART_REQ_DONE) return MCSART_OK;
    if (ended(r)) return fail(r->status, r->error);   // the cause, on the caller's thread
    return fail(MCSART_EHIP, sv->worker_error);
}

extern "C" int mcsart_server_drain(mcsart_server *sv) {
    if (!sv) return fail(MCSART_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(sv->mu);
    sv->cv_done.wait(lk, [&] {
        if (!sv->worker_error.empty()) return true;
        for (auto &kv : sv->reqs)
            if (!ended(kv.second)) return false;
        return true;
    });
    if (!sv->worker_error.empty()) return fail(MCSART_EHIP, sv->worker_error);
    return MCSART_OK;
}

extern "C" int mcsart_server_info(mcsart_server *sv, long long id, mcsart_request_info *info) {
    if (!sv || !info) return fail(MCSART_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return fail(MCSART_EINVAL, "unknown request id");
    const Request *r = it->second;
    std::memset(info, 0, sizeof(*info));
    info->state = r->state;
    info->status = r->status;
    info->slot = r->slot;
    info->bytes = r->bytes;
    info->tuning = r->tuning;
    const auto now = Clock::now();
    const bool admitted = r->slot >= 0;
    const bool fin = ended(r);
    info->queue_ms = ms_between(r->t_submit, admitted ? r->t_admit : (fin ? r->t_done : now));
    info->run_ms = admitted ? ms_between(r->t_admit, fin ? r->t_done : now) : 0.0;
    info->total_ms = ms_between(r->t_submit, fin ? r->t_done : now);
    info->admit_at_ms = admitted ? ms_between(sv->t0, r->t_admit) : 0.0;
    info->done_at_ms = fin ? ms_between(sv->t0, r->t_done) : 0.0;
    return MCSART_OK;
}

extern "C" int mcsart_server_release(mcsart_server *sv, long long id) {
    if (!sv) return fail(MCSART_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return fail(MCSART_EINVAL, "unknown request id");
    if (!ended(it->second)) return fail(MCSART_EBUSY, "request still queued or running");
    delete it->second;
    sv->reqs.erase(it);
    return MCSART_OK;
}

extern "C" int mcsart_server_memory(mcsart_server *sv, size_t *used, size_t *budget) {
    if (!sv) return fail(MCSART_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(sv->mu);
    if (used) *used = sv->used - idle_bytes(sv);   // held by admitted jobs
    if (budget) *budget = sv->budget;
    return MCSART_OK;
}

extern "C" int mcsart_server_destroy(mcsart_server *sv) {
    if (!sv) return MCSART_OK;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;   // queued requests still run: each slot exits once the queue is empty
        for (Slot *s : sv->slots) s->cv.notify_one();
        sv->cv_done.notify_all();
    }
    destroy_slots(sv);
    (void)hipSetDevice(sv->device);
    for (const Block &b : sv->idle) (void)hipFree(b.ptr);
    for (auto &kv : sv->reqs) delete kv.second;
    delete sv;
    return MCSART_OK;
}
