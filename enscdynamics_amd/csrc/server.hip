// Continuous-batching solve server (include/ensc_engine.h, ensc_server_*).
//
// Requests queue on the host.  Each pump admits queued jobs in arrival order
// while their device footprint fits the budget, starts them (their init
// pass runs inside the next step iteration of the live batch, then step 1 is
// switched in), advances every live job together for a number of step
// iterations, and retires the jobs that reached their end time: their
// results are copied to the host and their device memory is released.  A
// job's arithmetic never involves another job (per-job reductions in fixed
// order), so its result is the same alone or in any mix.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "engine_host.h"

namespace ensc {
ensc_status engine_new_public(const ensc_options* o, ensc_engine** out);
ensc_status engine_add_job(ensc_engine* E, const ensc_model_desc* m, uint32_t* slot);
ensc_status engine_start_jobs(ensc_engine* E, const std::vector<uint32_t>& slots);
ensc_status engine_ticks(ensc_engine* E, int64_t n);
ensc_status engine_poll(ensc_engine* E, std::vector<uint32_t>& finished);
ensc_status engine_job_scalars(ensc_engine* E, uint32_t s, double* out);
ensc_status engine_job_download(ensc_engine* E, uint32_t s, const char* f, double* out, int64_t count);
ensc_status engine_remove_job(ensc_engine* E, uint32_t s);
int64_t engine_job_bytes(ensc_engine* E, uint32_t s);
ensc_status engine_set_graph(ensc_engine* E, int32_t ticks);
}  // namespace ensc

using namespace ensc;

namespace {

// A deep copy of a submitted model, with its own ensc_model_desc view.
struct OwnedModel {
    ModelCopy c;
    std::vector<std::vector<ensc_bc_entry>> spc, vel;
    std::vector<ensc_step_def> steps;
    std::vector<ensc_amplitude> amps;
    ensc_model_desc d{};
    void bind() {
        spc.assign(c.steps.size(), {});
        vel.assign(c.steps.size(), {});
        steps.assign(c.steps.size(), ensc_step_def{});
        for (size_t s = 0; s < c.steps.size(); ++s) {
            const StepCopy& S = c.steps[s];
            auto mk = [](const BcCopy& b) {
                ensc_bc_entry e{};
                e.nodes = b.nodes.data();
                e.n_nodes = (uint32_t)b.nodes.size();
                e.d0 = b.d0;
                e.d1 = b.d1;
                e.cond = b.cond;
                e.value = b.value;
                e.amp = b.amp;
                return e;
            };
            for (const auto& b : S.spc) spc[s].push_back(mk(b));
            for (const auto& b : S.vel) vel[s].push_back(mk(b));
            ensc_step_def& t = steps[s];
            t.period = S.period;
            t.reset_spc = S.reset_spc;
            t.reset_vel = S.reset_vel;
            t.reset_dload = S.reset_dload;
            t.grav_on = S.grav_on;
            t.grav_value = S.grav_value;
            t.grav_amp = S.grav_amp;
            for (int q = 0; q < 3; ++q) t.grav_dir[q] = S.grav_dir[q];
            t.n_spc = (uint32_t)spc[s].size();
            t.spc = spc[s].data();
            t.n_vel = (uint32_t)vel[s].size();
            t.vel = vel[s].data();
        }
        const size_t na = c.amp_off.empty() ? 0 : c.amp_off.size() - 1;
        amps.assign(na, ensc_amplitude{});
        for (size_t i = 0; i < na; ++i) {
            amps[i].t = c.amp_t.data() + c.amp_off[i];
            amps[i].a = c.amp_a.data() + c.amp_off[i];
            amps[i].n = (uint32_t)(c.amp_off[i + 1] - c.amp_off[i]);
        }
        d = ensc_model_desc{};
        d.n_nodes = c.Nn;
        d.n_elems = c.Ne;
        d.x0 = c.x0.data();
        d.conn = c.conn.data();
        d.E = c.E;
        d.nu = c.nu;
        d.rho = c.rho;
        d.plastic = c.plastic;
        d.yield0 = c.yield0;
        d.hard = c.hard;
        d.c_dr = c.c_dr;
        d.c_cr = c.c_cr;
        d.factor = c.factor;
        d.cvisl = c.cvisl;
        d.cvisq = c.cvisq;
        d.total_time = c.total_time;
        d.v_init = c.v_init.data();
        d.n_steps = (uint32_t)steps.size();
        d.steps = steps.data();
        d.n_amps = (uint32_t)amps.size();
        d.amps = amps.empty() ? nullptr : amps.data();
    }
};

struct Request {
    uint64_t ticket = 0;
    std::unique_ptr<OwnedModel> model;
    int64_t footprint = 0;
    int32_t keep = 0;
    double t_submit = 0.0;
};

struct Record {
    double t_submit = 0, t_admit = -1, t_finish = -1;
    double scalars[ENSC_S_COUNT] = {0};
    std::vector<double> u;
    uint32_t n_nodes = 0;
    int32_t keep = 0;
    bool finished = false;
};

}  // namespace

struct ensc_server {
    ensc_engine* eng = nullptr;
    int precision = ENSC_FP64;
    int64_t budget = 0, used = 0;
    int32_t max_live = 0;
    std::chrono::steady_clock::time_point t0;
    uint64_t next_ticket = 1;
    std::deque<Request> queue;
    std::map<uint32_t, std::pair<uint64_t, int64_t>> live;  // slot -> (ticket, bytes)
    std::map<uint64_t, Record> rec;
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

// Device bytes a job will hold (matches build_job's allocations).
int64_t footprint(const ModelCopy& c, int precision) {
    const int64_t es = (int64_t)round_up(c.Ne, 256), nn = c.Nn;
    const int64_t r = precision == ENSC_MIXED ? 4 : 8;
    // conn, lconn, Gi, V0, sigma, QIA, plastic strain, corner forces
    const int64_t elem = 32 + 16 + 9 * r + r + 6 * r + 12 * r + r + 24 * r;
    // record, reactions, BC mask/values/amplitudes, 8-wide gather table,
    // staging lists (< 2 positions per node); mixed: u compensation and the
    // block-assembly partials and position lists
    const int64_t node = NODE_BYTES + 24 + 1 + 24 + 6 + 8 * 4 + 8 +
                         (precision == ENSC_MIXED ? 16 + 4 + 2 * (4 + 12) : 0) + 16;
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
    S->budget = hbm_budget_bytes > 0 ? hbm_budget_bytes : (int64_t)64 << 30;
    S->max_live = max_live_jobs > 0 ? max_live_jobs : 1024;
    S->t0 = std::chrono::steady_clock::now();
    *out = S.release();
    return ENSC_OK;
}

ensc_status ensc_server_destroy(ensc_server* S) {
    if (!S) return ENSC_OK;
    ensc_engine_destroy(S->eng);
    delete S;
    return ENSC_OK;
}

ensc_status ensc_server_set_graph(ensc_server* S, int32_t ticks) {
    if (!S) return fail(ENSC_ERR_ARG, "null server");
    return engine_set_graph(S->eng, ticks);
}

ensc_status ensc_server_submit(ensc_server* S, const ensc_model_desc* model, int32_t keep_result, double t_arrival,
                               uint64_t* ticket) {
    if (!S || !ticket) return fail(ENSC_ERR_ARG, "null argument");
    Request r;
    r.model.reset(new OwnedModel());
    std::string err;
    if (!copy_model(model, r.model->c, err)) return fail(ENSC_ERR_ARG, err);
    r.model->bind();
    r.footprint = footprint(r.model->c, S->precision);
    if (r.footprint > S->budget) return fail(ENSC_ERR_MEM, "job larger than the server's device budget");
    r.ticket = S->next_ticket++;
    r.keep = keep_result;
    r.t_submit = t_arrival >= 0.0 ? t_arrival : S->now();
    Record& rc = S->rec[r.ticket];
    rc.t_submit = r.t_submit;
    rc.keep = keep_result;
    *ticket = r.ticket;
    S->queue.push_back(std::move(r));
    return ENSC_OK;
}

ensc_status ensc_server_pump(ensc_server* S, int32_t ticks, uint64_t* done, uint32_t cap, uint32_t* n_done) {
    if (!S || !n_done) return fail(ENSC_ERR_ARG, "null argument");
    *n_done = 0;
    ensc_status st;
    // 1. admission by footprint, arrival order
    std::vector<uint32_t> started;
    while (!S->queue.empty() && (int32_t)S->live.size() < S->max_live &&
           S->used + S->queue.front().footprint <= S->budget) {
        Request r = std::move(S->queue.front());
        S->queue.pop_front();
        uint32_t slot = 0;
        if ((st = engine_add_job(S->eng, &r.model->d, &slot))) return st;
        const int64_t bytes = engine_job_bytes(S->eng, slot);
        S->used += bytes;
        S->live[slot] = {r.ticket, bytes};
        S->rec[r.ticket].t_admit = S->now();
        S->rec[r.ticket].n_nodes = r.model->c.Nn;
        started.push_back(slot);
    }
    // 2. new jobs: init pass inside the next step iteration, then step 1
    if (!started.empty() && (st = engine_start_jobs(S->eng, started))) return st;
    // 3. advance every live job together
    if (!S->live.empty() && ticks > 0 && (st = engine_ticks(S->eng, ticks))) return st;
    // 4. retire finished jobs: results to the host, device memory released
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
        if ((st = engine_remove_job(S->eng, slot))) return st;
        rc.t_finish = S->now();
        rc.finished = true;
        S->used -= it->second.second;
        S->live.erase(it);
        if (done && *n_done < cap) done[*n_done] = t;
        ++*n_done;
    }
    return ENSC_OK;
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
