// Reconstruction server: continuous batching of McSART jobs on one GPU.
//
// No counterpart in the reference, whose driver reconstructs one scan per MATLAB call
// (src/entry_AUMISART.cu:193-207).  Requests queue in arrival order; a scheduler thread admits a
// request as soon as a slot (a stream with its own context) is free and the request's device
// footprint (mcsart_job_bytes) fits in the remaining HBM budget.  An admitted request becomes a
// device-resident job: inputs uploaded once, every sweep enqueued on the slot's stream as a
// replayed HIP graph, an event recorded behind the last sweep.  Jobs on different slots overlap
// on the device (their kernels fill each other's idle compute units and hide each other's launch
// gaps).  When a slot's event completes, the scheduler copies the results into the caller's
// buffers, frees the job and admits the next request, so memory and slots are reused while other
// jobs keep running.  A job's kernels, launch shapes and their order do not depend on what else
// runs, so a request's results are the same alone and in any mix.
#include <algorithm>
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

struct Request {
    long long id = 0;
    mcsart_request in{};
    std::vector<float> angles, volumes, flows;   // small per-view arrays are copied at submit
    std::vector<int> op_iter;
    size_t bytes = 0;
    int state = MCSART_REQ_QUEUED;
    int status = MCSART_OK;
    int slot = -1;
    int bypassed = 0;   // times a later request was admitted ahead of this one
    mcsart_tuning tuning{0, 0, 0, 0};
    mcsart_job *job = nullptr;
    Clock::time_point t_submit, t_admit, t_done;
};

struct Slot {
    mcsart_ctx *ctx = nullptr;
    hipEvent_t done = nullptr;
    Request *req = nullptr;
};

}  // namespace

struct mcsart_server {
    int device = 0;
    mcsart_server_opts opts{};
    size_t budget = 0, used = 0;
    std::vector<Slot> slots;
    std::deque<Request *> queue;
    std::map<long long, Request *> reqs;
    long long next_id = 1;
    std::mutex mu;
    std::condition_variable cv_work;   // new submissions / shutdown
    std::condition_variable cv_done;   // a request finished
    bool stop = false;
    std::thread worker;
    std::string worker_error;
    Clock::time_point t0 = Clock::now();   // origin of the requests' admit/done times
};

namespace {

// Upload a request's inputs into a fresh job on slot `s` and enqueue its whole schedule.
int start_request(mcsart_server *sv, Slot &s, Request *r) {
    const mcsart_request &q = r->in;
    mcsart_job *j = nullptr;
    int rc = mcsart_job_create(s.ctx, &q.geom, q.n_views, r->angles.data(), r->volumes.data(),
                               r->flows.data(), &q.opts, q.n_iter, r->op_iter.data(), &j);
    if (rc) return rc;
    r->job = j;
    if ((rc = mcsart_job_set_projections(j, q.h_proj))) return rc;
    if ((rc = mcsart_job_set_image(j, q.h_img))) return rc;
    if ((rc = mcsart_job_set_fields(j, q.h_fields))) return rc;   // NULL components: zero
    if (sv->opts.autotune) {
        if ((rc = mcsart_job_autotune(j, &r->tuning))) return rc;
    } else {
        mcsart_tuning t{sv->opts.proj_split, sv->opts.back_zslab, sv->opts.back_stage,
                        sv->opts.proj_rows, sv->opts.back_zslab_motion};
        if ((rc = mcsart_job_set_tuning(j, &t))) return rc;
        r->tuning = t;
    }
    if ((rc = mcsart_job_run(j))) return rc;
    MCSART_HIP_TRY(hipEventRecord(s.done, s.ctx->stream));
    return MCSART_OK;
}

// Results into the caller's buffers, then the job's memory is released.
int finish_request(Request *r) {
    const mcsart_request &q = r->in;
    int rc = MCSART_OK;
    if (q.out_img) rc = mcsart_job_get_image(r->job, q.out_img);
    if (!rc && q.out_norm) rc = mcsart_job_get_norms(r->job, q.out_norm);
    bool any = false;
    for (int c = 0; c < 6; ++c) any |= q.out_fields[c] != nullptr;
    if (!rc && any) rc = mcsart_job_get_fields(r->job, q.out_fields);
    mcsart_job_destroy(r->job);
    r->job = nullptr;
    return rc;
}

// First request in arrival order that fits the free budget.  A request that fits nowhere yet
// is bypassed at most 2*slots times; after that nothing is admitted ahead of it (no starvation).
Request *pick(mcsart_server *sv) {
    const size_t free_bytes = sv->budget - sv->used;
    for (size_t k = 0; k < sv->queue.size(); ++k) {
        Request *r = sv->queue[k];
        if (r->bytes <= free_bytes) {
            for (size_t m = 0; m < k; ++m) sv->queue[m]->bypassed++;
            sv->queue.erase(sv->queue.begin() + (long)k);
            return r;
        }
        if (r->bypassed >= 2 * (int)sv->slots.size()) return nullptr;
    }
    return nullptr;
}

void worker_main(mcsart_server *sv) {
    if (hipSetDevice(sv->device) != hipSuccess) {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->worker_error = "hipSetDevice failed in the scheduler thread";
        sv->stop = true;
        sv->cv_done.notify_all();
        return;
    }
    std::unique_lock<std::mutex> lk(sv->mu);
    while (true) {
        bool progressed = false;
        // 1. completions
        for (auto &s : sv->slots) {
            if (!s.req) continue;
            const hipError_t e = hipEventQuery(s.done);
            if (e == hipErrorNotReady) continue;
            Request *r = s.req;
            lk.unlock();
            int rc = (e == hipSuccess) ? finish_request(r)
                                       : fail(MCSART_EHIP, std::string("job failed: ") +
                                                               hipGetErrorString(e));
            if (rc && r->job) {
                mcsart_job_destroy(r->job);
                r->job = nullptr;
            }
            lk.lock();
            r->t_done = Clock::now();
            r->status = rc;
            r->state = rc ? MCSART_REQ_FAILED : MCSART_REQ_DONE;
            sv->used -= r->bytes;
            s.req = nullptr;
            progressed = true;
            sv->cv_done.notify_all();
        }
        // 2. admissions
        for (auto &s : sv->slots) {
            if (s.req) continue;
            Request *r = pick(sv);
            if (!r) break;
            r->state = MCSART_REQ_RUNNING;
            r->slot = (int)(&s - sv->slots.data());
            r->t_admit = Clock::now();
            sv->used += r->bytes;
            s.req = r;
            lk.unlock();
            const int rc = start_request(sv, s, r);
            lk.lock();
            if (rc) {   // the request fails alone; the server keeps going
                if (r->job) {
                    mcsart_job_destroy(r->job);
                    r->job = nullptr;
                }
                r->status = rc;
                r->state = MCSART_REQ_FAILED;
                r->t_done = Clock::now();
                sv->used -= r->bytes;
                s.req = nullptr;
                sv->cv_done.notify_all();
            }
            progressed = true;
        }
        bool busy = false;
        for (auto &s : sv->slots) busy |= s.req != nullptr;
        if (sv->stop && !busy && sv->queue.empty()) break;
        if (!progressed) {
            // running jobs: poll their events every 100 us; idle: sleep until a submission
            if (busy) {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(100));
                lk.lock();
            } else {
                sv->cv_work.wait(lk, [&] { return sv->stop || !sv->queue.empty(); });
            }
        }
    }
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
        Slot s;
        int rc = mcsart_ctx_create(&s.ctx, device);
        if (!rc && hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess)
            rc = fail(MCSART_EHIP, "hipEventCreate failed");
        if (rc) {
            if (s.ctx) mcsart_ctx_destroy(s.ctx);
            for (auto &x : sv->slots) {
                (void)hipEventDestroy(x.done);
                mcsart_ctx_destroy(x.ctx);
            }
            delete sv;
            return rc;
        }
        sv->slots.push_back(s);
    }
    sv->worker = std::thread(worker_main, sv);
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
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sv->stop) {
            delete r;
            return fail(MCSART_EBUSY, "server is shutting down");
        }
        r->id = sv->next_id++;
        sv->reqs[r->id] = r;
        sv->queue.push_back(r);
        if (id) *id = r->id;
    }
    sv->cv_work.notify_all();
    return MCSART_OK;
}

extern "C" int mcsart_server_wait(mcsart_server *sv, long long id, double timeout_s) {
    if (!sv) return fail(MCSART_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return fail(MCSART_EINVAL, "unknown request id");
    Request *r = it->second;
    auto done = [&] {
        return r->state == MCSART_REQ_DONE || r->state == MCSART_REQ_FAILED || sv->stop;
    };
    if (timeout_s < 0) {
        sv->cv_done.wait(lk, done);
    } else if (!sv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) {
        return fail(MCSART_EBUSY, "request " + std::to_string(id) + " still running");
    }
    if (r->state == MCSART_REQ_FAILED) return r->status;
    if (r->state != MCSART_REQ_DONE)
        return fail(MCSART_EHIP, sv->worker_error.empty() ? "server stopped" : sv->worker_error);
    return MCSART_OK;
}

extern "C" int mcsart_server_drain(mcsart_server *sv) {
    if (!sv) return fail(MCSART_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(sv->mu);
    sv->cv_done.wait(lk, [&] {
        if (sv->stop && !sv->worker_error.empty()) return true;
        for (auto &kv : sv->reqs)
            if (kv.second->state == MCSART_REQ_QUEUED || kv.second->state == MCSART_REQ_RUNNING)
                return false;
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
    const bool admitted = r->state != MCSART_REQ_QUEUED;
    const bool ended = r->state == MCSART_REQ_DONE || r->state == MCSART_REQ_FAILED;
    info->queue_ms = ms_between(r->t_submit, admitted ? r->t_admit : now);
    info->run_ms = admitted ? ms_between(r->t_admit, ended ? r->t_done : now) : 0.0;
    info->total_ms = ms_between(r->t_submit, ended ? r->t_done : now);
    info->admit_at_ms = admitted ? ms_between(sv->t0, r->t_admit) : 0.0;
    info->done_at_ms = ended ? ms_between(sv->t0, r->t_done) : 0.0;
    return MCSART_OK;
}

extern "C" int mcsart_server_release(mcsart_server *sv, long long id) {
    if (!sv) return fail(MCSART_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return fail(MCSART_EINVAL, "unknown request id");
    if (it->second->state != MCSART_REQ_DONE && it->second->state != MCSART_REQ_FAILED)
        return fail(MCSART_EBUSY, "request still queued or running");
    delete it->second;
    sv->reqs.erase(it);
    return MCSART_OK;
}

extern "C" int mcsart_server_memory(mcsart_server *sv, size_t *used, size_t *budget) {
    if (!sv) return fail(MCSART_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(sv->mu);
    if (used) *used = sv->used;
    if (budget) *budget = sv->budget;
    return MCSART_OK;
}

extern "C" int mcsart_server_destroy(mcsart_server *sv) {
    if (!sv) return MCSART_OK;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;   // queued requests still run: the scheduler exits once all are done
    }
    sv->cv_work.notify_all();
    if (sv->worker.joinable()) sv->worker.join();
    for (auto &s : sv->slots) {
        (void)hipEventDestroy(s.done);
        mcsart_ctx_destroy(s.ctx);
    }
    for (auto &kv : sv->reqs) delete kv.second;
    delete sv;
    return MCSART_OK;
}
