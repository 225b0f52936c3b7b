// Continuous-batching solve server (include/ensc_engine.h, ensc_server_*).
//
// Requests queue on the host.  Worker threads prepare queued jobs (setup
// arithmetic, renumbering, gather and staging tables: prepare_job) while the
// GPU steps the live ones.  Each pump admits prepared jobs in arrival order
// while their device footprint fits the budget (upload_job), starts them
// (their init pass runs inside the next step iteration of the live batch,
// then step 1 is switched in), advances every live job together for a
// number of step iterations, and retires the jobs that reached their end
// time: their results are copied to the host, and their device memory is
// released in groups (pump, step 4).  A job's arithmetic never involves another job (per-job
// reductions in fixed order), so its result is the same alone or in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "engine_host.h"

namespace ensc {
ensc_status engine_new_public(const ensc_options* o, ensc_engine** out);
ensc_status engine_prepare_job(const ensc_engine* E, ModelCopy&& model, std::unique_ptr<Job>& out);
ensc_status engine_add_prepared(ensc_engine* E, std::unique_ptr<Job> J, uint32_t* slot);
ensc_status engine_start_jobs(ensc_engine* E, const std::vector<uint32_t>& slots);
ensc_status engine_ticks(ensc_engine* E, int64_t n);
ensc_status engine_poll(ensc_engine* E, std::vector<uint32_t>& finished);
ensc_status engine_job_scalars(ensc_engine* E, uint32_t s, double* out);
ensc_status engine_job_download(ensc_engine* E, uint32_t s, const char* f, double* out, int64_t count);
ensc_status engine_remove_job(ensc_engine* E, uint32_t s);
int64_t engine_job_bytes(ensc_engine* E, uint32_t s);
uint32_t engine_job_blocks(ensc_engine* E, uint32_t s);
ensc_status engine_set_graph(ensc_engine* E, int32_t ticks);
ensc_status engine_bind_device(ensc_engine* E);
}  // namespace ensc

using namespace ensc;

namespace {

enum PrepState : int { PREP_QUEUED = 0, PREP_RUNNING = 1, PREP_READY = 2, PREP_FAILED = 3 };

struct Request {
    uint64_t ticket = 0;
    ModelCopy model;           // moved into the job by the worker
    std::unique_ptr<Job> job;  // prepared host image
    int64_t footprint = 0;
    int32_t keep = 0;
    double t_submit = 0.0;
    uint32_t n_nodes = 0;
    int state = PREP_QUEUED;  // guarded by ensc_server::mu
    std::string err;
};

struct Record {
    double t_submit = 0, t_admit = -1, t_finish = -1;
    double scalars[ENSC_S_COUNT] = {0};
    std::vector<double> u;
    uint32_t n_nodes = 0;
    int32_t keep = 0;
    bool finished = false;
    bool failed = false;
    std::string err;
};

}  // namespace

struct ensc_server {
    ensc_engine* eng = nullptr;
    int precision = ENSC_FP64, fp64_rule = 0;
    int64_t budget = 0, used = 0;
    int32_t max_live = 0;
    std::chrono::steady_clock::time_point t0;
    uint64_t next_ticket = 1;
    std::deque<std::shared_ptr<Request>> queue;  // admission order
    std::map<uint32_t, std::pair<uint64_t, int64_t>> live;  // slot -> (ticket, bytes)
    // finished jobs whose results are out but which are still in the engine
    // (slot, bytes): removed together, so the launch layout changes rarely
    std::vector<std::pair<uint32_t, int64_t>> retired;
    std::map<uint64_t, Record> rec;
    // preparation workers
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::shared_ptr<Request>> todo;  // preparation order (= arrival order)
    std::vector<std::thread> workers;
    bool stop = false;
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    void work() {
        for (;;) {
            std::shared_ptr<Request> r;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || !todo.empty(); });
                if (stop) return;
                r = todo.front();
                todo.pop_front();
                r->state = PREP_RUNNING;
            }
            std::unique_ptr<Job> J;
            const ensc_status st = guarded([&] { return engine_prepare_job(eng, std::move(r->model), J); });
            const std::string err = st ? std::string(ensc_last_error()) : std::string();
            {
                std::lock_guard<std::mutex> lk(mu);
                r->job = std::move(J);
                r->err = err;
                r->state = st ? PREP_FAILED : PREP_READY;
            }
            cv.notify_all();
        }
    }
};

namespace {

// Device bytes a job will hold (matches build_job's allocations).
int64_t footprint(const ModelCopy& c, int engine_precision, int fp64_rule) {
    const int64_t es = (int64_t)round_up(c.Ne, 256), nn = c.Nn;
    const int precision = job_precision(c, engine_precision, fp64_rule);
    const bool mixed = precision == ENSC_MIXED;
    const int64_t r = mixed ? 4 : 8;
    // conn, lconn, Gi and V0 (FP64 engine), sigma, QIA, plastic strain,
    // corner forces
    const int64_t elem = 32 + 16 + (mixed ? 0 : 10 * r) + 6 * r + 12 * r + r + 24 * r;
    // record, reactions, BC mask/values/amplitudes, 8-wide gather table,
    // staging lists (< 2 positions per node); mixed: u compensation, 1/m
    // and the block-assembly partials and position lists
    const int64_t node = NODE_BYTES + 24 + 1 + 24 + 6 + 8 * 4 + 8 +
                         (precision == ENSC_MIXED ? 4 + 8 + 4 + 2 * (4 + 12) : 0) + 16;
    return es * elem + nn * node + 64 * 1024;
}

}  // namespace

extern "C" {

ensc_status ensc_server_create(const ensc_options* opts, int64_t hbm_budget_bytes, int32_t max_live_jobs,
                               ensc_server** out) {
    if (!out) return fail(ENSC_ERR_ARG, "out is null");
    *out = nullptr;
    std::unique_ptr<ensc_server> S(new ensc_server());
    ensc_status st = engine_new_public(opts, &S->eng);
    if (st) return st;
    S->precision = opts ? opts->precision : ENSC_FP64;
    S->fp64_rule = opts ? opts->fp64_jobs : 0;
    S->budget = hbm_budget_bytes > 0 ? hbm_budget_bytes : (int64_t)64 << 30;
    S->max_live = max_live_jobs > 0 ? max_live_jobs : 1024;
    S->t0 = std::chrono::steady_clock::now();
    const unsigned hc = std::max(1u, std::thread::hardware_concurrency());
    const unsigned nw = std::min(4u, hc);
    ensc_server* raw = S.get();
    for (unsigned i = 0; i < nw; ++i) S->workers.emplace_back([raw] { raw->work(); });
    *out = S.release();
    return ENSC_OK;
}

ensc_status ensc_server_destroy(ensc_server* S) {
    if (!S) return ENSC_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv.notify_all();
    for (auto& t : S->workers) t.join();
    ensc_engine_destroy(S->eng);
    delete S;
    return ENSC_OK;
}

ensc_status ensc_server_set_graph(ensc_server* S, int32_t ticks) {
    if (!S) return fail(ENSC_ERR_ARG, "null server");
    ensc_status st = engine_bind_device(S->eng);
    if (st) return st;
    return engine_set_graph(S->eng, ticks);
}

ensc_status ensc_server_submit(ensc_server* S, const ensc_model_desc* model, int32_t keep_result, double t_arrival,
                               uint64_t* ticket) {
    if (!S || !ticket) return fail(ENSC_ERR_ARG, "null argument");
    return guarded([&]() -> ensc_status {
        std::shared_ptr<Request> r(new Request());
        std::string err;
        if (!copy_model(model, r->model, err)) return fail(ENSC_ERR_ARG, err);
        r->footprint = footprint(r->model, S->precision, S->fp64_rule);
        if (r->footprint > S->budget) return fail(ENSC_ERR_MEM, "job larger than the server's device budget");
        r->ticket = S->next_ticket++;
        r->keep = keep_result;
        r->n_nodes = r->model.Nn;
        r->t_submit = t_arrival >= 0.0 ? t_arrival : S->now();
        Record& rc = S->rec[r->ticket];
        rc.t_submit = r->t_submit;
        rc.keep = keep_result;
        rc.n_nodes = r->n_nodes;
        *ticket = r->ticket;
        S->queue.push_back(r);
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->todo.push_back(r);
        }
        S->cv.notify_all();
        return ENSC_OK;
    });
}

ensc_status ensc_server_pump(ensc_server* S, int32_t ticks, uint64_t* done, uint32_t cap, uint32_t* n_done) {
    if (!S || !n_done) return fail(ENSC_ERR_ARG, "null argument");
    return guarded([&]() -> ensc_status {
        *n_done = 0;
        ensc_status st;
        if ((st = engine_bind_device(S->eng))) return st;
        auto report = [&](uint64_t t) {
            if (done && *n_done < cap) done[*n_done] = t;
            ++*n_done;
        };
        auto remove_retired = [&]() -> ensc_status {
            for (const auto& z : S->retired) {
                const ensc_status s2 = engine_remove_job(S->eng, z.first);
                if (s2) return s2;
                S->used -= z.second;
            }
            S->retired.clear();
            return ENSC_OK;
        };
        // 1. admission of prepared jobs by footprint, arrival order; with nothing
        //    on the device, wait for the next job's preparation
        std::vector<uint32_t> started;
        while (!S->queue.empty() && (int32_t)S->live.size() < S->max_live) {
            const std::shared_ptr<Request> r = S->queue.front();
            int state;
            {
                std::unique_lock<std::mutex> lk(S->mu);
                if (r->state < PREP_READY && S->live.empty() && started.empty())
                    S->cv.wait(lk, [&] { return r->state >= PREP_READY; });
                state = r->state;
            }
            if (state < PREP_READY) break;
            if (state == PREP_FAILED) {
                Record& rc = S->rec[r->ticket];
                rc.failed = rc.finished = true;
                rc.err = r->err;
                rc.t_admit = rc.t_finish = S->now();
                S->queue.pop_front();
                report(r->ticket);
                continue;
            }
            if (S->used + r->footprint > S->budget && !S->retired.empty() && (st = remove_retired())) return st;
            if (S->used + r->footprint > S->budget) break;
            S->queue.pop_front();
            uint32_t slot = 0;
            if (engine_add_prepared(S->eng, std::move(r->job), &slot)) {
                // could not be placed on the device (an allocation or upload
                // failed): retired with its reason, like a failed set-up
                Record& rc = S->rec[r->ticket];
                rc.failed = rc.finished = true;
                rc.err = ensc_last_error();
                rc.t_admit = rc.t_finish = S->now();
                report(r->ticket);
                continue;
            }
            const int64_t bytes = engine_job_bytes(S->eng, slot);
            S->used += bytes;
            S->live[slot] = {r->ticket, bytes};
            S->rec[r->ticket].t_admit = S->now();
            started.push_back(slot);
        }
        // 2. new jobs: init pass inside the next step iteration, then step 1
        if (!started.empty() && (st = engine_start_jobs(S->eng, started))) return st;
        // 3. advance every live job together
        if (!S->live.empty() && ticks > 0 && (st = engine_ticks(S->eng, ticks))) return st;
        // 4. jobs at the end of a *Step go on with the next one (engine_poll);
        //    retire finished jobs: results to the host.  Their device memory is
        //    released once the retired jobs hold a quarter of the element blocks
        //    still in the launch, when nothing live is left, or when an
        //    admission needs the room (a finished job's blocks exit at once, so
        //    keeping it in the layout costs little, while every layout change
        //    rebuilds the block maps and the step graph).
        std::vector<uint32_t> fin;
        if ((st = engine_poll(S->eng, fin))) return st;
        for (uint32_t slot : fin) {
            auto it = S->live.find(slot);
            if (it == S->live.end()) continue;
            const uint64_t t = it->second.first;
            Record& rc = S->rec[t];
            if ((st = engine_job_scalars(S->eng, slot, rc.scalars))) return st;
            if (rc.keep) {
                rc.u.assign(3 * (size_t)rc.n_nodes, 0.0);
                if ((st = engine_job_download(S->eng, slot, "u", rc.u.data(), (int64_t)rc.u.size()))) return st;
            }
            rc.t_finish = S->now();
            rc.finished = true;
            S->retired.emplace_back(slot, it->second.second);
            S->live.erase(it);
            report(t);
        }
        if (!S->retired.empty()) {
            uint64_t zb = 0, lb = 0;
            for (const auto& z : S->retired) zb += engine_job_blocks(S->eng, z.first);
            for (const auto& l : S->live) lb += engine_job_blocks(S->eng, l.first);
            if (S->live.empty() || zb * 4 > zb + lb) return remove_retired();
        }
        return ENSC_OK;
    });
}

ensc_status ensc_server_result(ensc_server* S, uint64_t ticket, double* scalars, double* u, int64_t u_count,
                               double* times /* submit, admit, finish (s) */) {
    if (!S) return fail(ENSC_ERR_ARG, "null server");
    auto it = S->rec.find(ticket);
    if (it == S->rec.end()) return fail(ENSC_ERR_ARG, "unknown ticket");
    const Record& rc = it->second;
    if (times) {
        times[0] = rc.t_submit;
        times[1] = rc.t_admit;
        times[2] = rc.t_finish;
    }
    if (!rc.finished) return fail(ENSC_ERR_STATE, "job not finished");
    if (rc.failed) return fail(ENSC_ERR_ARG, "job could not be set up: " + rc.err);
    if (scalars) std::memcpy(scalars, rc.scalars, sizeof(rc.scalars));
    if (u) {
        if (!rc.keep) return fail(ENSC_ERR_STATE, "result fields were not kept for this ticket");
        if (u_count < (int64_t)rc.u.size()) return fail(ENSC_ERR_ARG, "buffer too small");
        std::memcpy(u, rc.u.data(), rc.u.size() * sizeof(double));
    }
    return ENSC_OK;
}

ensc_status ensc_server_status(ensc_server* S, uint32_t* queued, uint32_t* live, int64_t* bytes_used) {
    if (!S) return fail(ENSC_ERR_ARG, "null server");
    if (queued) *queued = (uint32_t)S->queue.size();
    if (live) *live = (uint32_t)S->live.size();
    if (bytes_used) *bytes_used = S->used;
    return ENSC_OK;
}

double ensc_server_clock(ensc_server* S) { return S ? S->now() : 0.0; }

}  // extern "C"
