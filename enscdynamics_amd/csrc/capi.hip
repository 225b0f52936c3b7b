// C ABI of the step engine (include/ensc_engine.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "engine_host.h"
#include "kernels.h"

using namespace ensc;

#define HIPCHK(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return fail(ENSC_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

typedef void (*KernelFn)(const JobDev*, const uint32_t*);

struct ensc_engine {
    int device = 0, precision = ENSC_FP64, blk_e = 256, blk_n = 256;
    hipStream_t stream = nullptr;
    std::vector<std::unique_ptr<Job>> jobs;
    DeviceArena arena;
    JobDev* d_jobs = nullptr;
    uint32_t* d_b2j_e = nullptr;
    uint32_t* d_b2j_n = nullptr;
    uint32_t nblk_e = 0, nblk_n = 0;
    Ctl* d_ctl = nullptr;
    Ctl* h_ctl = nullptr;  // pinned host mirror
    double* d_part_e = nullptr;
    double* d_part_n = nullptr;
    KernelFn ke = nullptr, kn = nullptr;
    bool timing = false;
    std::vector<hipEvent_t> evs;
    double elem_ms = 0.0, node_ms = 0.0;
    int64_t launches = 0;
    int64_t pending_ticks = 0;
};

namespace {

template <bool M>
KernelFn pick_e(int blk) {
    switch (blk) {
        case 64: return k_element<M, 64>;
        case 128: return k_element<M, 128>;
        default: return k_element<M, 256>;
    }
}
template <bool M>
KernelFn pick_n(int blk) {
    switch (blk) {
        case 64: return k_node<M, 64>;
        case 128: return k_node<M, 128>;
        default: return k_node<M, 256>;
    }
}

ensc_status upload_jobs(ensc_engine* E) {
    std::vector<JobDev> h(E->jobs.size());
    for (size_t j = 0; j < E->jobs.size(); ++j) h[j] = E->jobs[j]->dev;
    HIPCHK(hipMemcpyAsync(E->d_jobs, h.data(), h.size() * sizeof(JobDev), hipMemcpyHostToDevice, E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
    return ENSC_OK;
}

ensc_status read_ctl(ensc_engine* E) {
    HIPCHK(hipMemcpyAsync(E->h_ctl, E->d_ctl, E->jobs.size() * sizeof(Ctl), hipMemcpyDeviceToHost, E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
    return ENSC_OK;
}

ensc_status write_ctl(ensc_engine* E) {
    HIPCHK(hipMemcpyAsync(E->d_ctl, E->h_ctl, E->jobs.size() * sizeof(Ctl), hipMemcpyHostToDevice, E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
    return ENSC_OK;
}

ensc_status enqueue_ticks(ensc_engine* E, int64_t n) {
    if (E->timing) {
        const size_t need = 2 * (size_t)n + 1;
        while (E->evs.size() < need) {
            hipEvent_t ev;
            HIPCHK(hipEventCreate(&ev));
            E->evs.push_back(ev);
        }
        HIPCHK(hipEventRecord(E->evs[0], E->stream));
    }
    for (int64_t i = 0; i < n; ++i) {
        hipLaunchKernelGGL(E->ke, dim3(E->nblk_e), dim3(E->blk_e), 0, E->stream, E->d_jobs, E->d_b2j_e);
        if (E->timing) HIPCHK(hipEventRecord(E->evs[2 * i + 1], E->stream));
        hipLaunchKernelGGL(E->kn, dim3(E->nblk_n), dim3(E->blk_n), 0, E->stream, E->d_jobs, E->d_b2j_n);
        if (E->timing) HIPCHK(hipEventRecord(E->evs[2 * i + 2], E->stream));
    }
    HIPCHK(hipGetLastError());
    E->pending_ticks = E->timing ? n : 0;
    return ENSC_OK;
}

ensc_status finish_timing(ensc_engine* E) {
    if (!E->timing || E->pending_ticks == 0) return ENSC_OK;
    double se = 0.0, sn = 0.0;
    for (int64_t i = 0; i < E->pending_ticks; ++i) {
        float a = 0.f, b = 0.f;
        HIPCHK(hipEventElapsedTime(&a, E->evs[2 * i], E->evs[2 * i + 1]));
        HIPCHK(hipEventElapsedTime(&b, E->evs[2 * i + 1], E->evs[2 * i + 2]));
        se += a;
        sn += b;
    }
    E->elem_ms = se / E->pending_ticks;
    E->node_ms = sn / E->pending_ticks;
    E->launches = E->pending_ticks;
    E->pending_ticks = 0;
    return ENSC_OK;
}

ensc_status sync(ensc_engine* E) {
    HIPCHK(hipStreamSynchronize(E->stream));
    return finish_timing(E);
}

bool job_done(const Ctl& c) { return !(c.time < c.step_end) || c.nsteps >= c.step_limit || c.err != 0; }

}  // namespace

extern "C" {

const char* ensc_last_error(void) { return g_last_error.c_str(); }

int32_t ensc_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

ensc_status ensc_engine_create(const ensc_model_desc* model, const ensc_options* opts, ensc_engine** out) {
    if (!out) return fail(ENSC_ERR_ARG, "out is null");
    *out = nullptr;
    ensc_options o{};
    if (opts) o = *opts;
    if (o.precision != ENSC_FP64 && o.precision != ENSC_MIXED) return fail(ENSC_ERR_ARG, "unknown precision");
    if (ensc_device_count() <= 0) return fail(ENSC_ERR_NODEV, "no HIP device visible");
    std::unique_ptr<ensc_engine> E(new ensc_engine());
    E->device = o.device;
    E->precision = o.precision;
    auto okblk = [](int b) { return b == 64 || b == 128 || b == 256; };
    E->blk_e = okblk(o.block_elem) ? o.block_elem : 256;
    E->blk_n = okblk(o.block_node) ? o.block_node : 256;
    HIPCHK(hipSetDevice(E->device));
    HIPCHK(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
    std::unique_ptr<Job> J(new Job());
    std::string err;
    if (!copy_model(model, J->model, err)) return fail(ENSC_ERR_ARG, err);
    ensc_status st = build_job(*J, E->precision, o.reorder, E->blk_e, E->blk_n);
    if (st) return st;
    // probes and history
    if (o.n_probe > MAX_PROBE) return fail(ENSC_ERR_ARG, "at most 8 probe nodes");
    for (uint32_t i = 0; i < o.n_probe; ++i) {
        if (o.probe_nodes[i] >= J->model.Nn) return fail(ENSC_ERR_ARG, "probe node out of range");
        J->dev.probe[i] = J->n_old2new[o.probe_nodes[i]];
    }
    J->dev.n_probe = (int32_t)o.n_probe;
    if (o.hist_cap > 0) {
        const size_t bytes = (size_t)o.hist_cap * (ENSC_H_BASE + 3 * o.n_probe) * sizeof(double);
        J->d_hist = static_cast<double*>(J->arena.alloc(bytes));
        if (!J->d_hist) return fail(ENSC_ERR_MEM, "history allocation failed");
        HIPCHK(hipMemset(J->d_hist, 0, bytes));
        J->dev.hist = J->d_hist;
        J->dev.hist_cap = o.hist_cap;
        J->hist_cap = o.hist_cap;
    }
    E->jobs.push_back(std::move(J));
    // engine-wide arrays: ctl, partials, block maps, job table
    const size_t nj = E->jobs.size();
    E->d_ctl = static_cast<Ctl*>(E->arena.alloc(nj * sizeof(Ctl)));
    E->d_jobs = static_cast<JobDev*>(E->arena.alloc(nj * sizeof(JobDev)));
    if (!E->d_ctl || !E->d_jobs) return fail(ENSC_ERR_MEM, "engine allocation failed");
    HIPCHK(hipHostMalloc((void**)&E->h_ctl, nj * sizeof(Ctl), hipHostMallocDefault));
    std::vector<uint32_t> b2e, b2n;
    for (size_t j = 0; j < nj; ++j) {
        JobDev& D = E->jobs[j]->dev;
        D.eblk0 = (uint32_t)b2e.size();
        D.nblk0 = (uint32_t)b2n.size();
        for (uint32_t b = 0; b < D.nblk_e; ++b) b2e.push_back((uint32_t)j);
        for (uint32_t b = 0; b < D.nblk_n; ++b) b2n.push_back((uint32_t)j);
    }
    E->nblk_e = (uint32_t)b2e.size();
    E->nblk_n = (uint32_t)b2n.size();
    E->d_b2j_e = static_cast<uint32_t*>(E->arena.alloc(b2e.size() * 4));
    E->d_b2j_n = static_cast<uint32_t*>(E->arena.alloc(b2n.size() * 4));
    E->d_part_e = static_cast<double*>(E->arena.alloc((size_t)E->nblk_e * 8 * sizeof(double)));
    E->d_part_n = static_cast<double*>(E->arena.alloc((size_t)E->nblk_n * 8 * sizeof(double)));
    if (!E->d_b2j_e || !E->d_b2j_n || !E->d_part_e || !E->d_part_n) return fail(ENSC_ERR_MEM, "engine allocation failed");
    HIPCHK(hipMemcpy(E->d_b2j_e, b2e.data(), b2e.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(E->d_b2j_n, b2n.data(), b2n.size() * 4, hipMemcpyHostToDevice));
    for (size_t j = 0; j < nj; ++j) {
        Job& Jb = *E->jobs[j];
        JobDev& D = Jb.dev;
        D.ctl = E->d_ctl + j;
        D.part_e = E->d_part_e + (size_t)D.eblk0 * 8;
        D.part_n = E->d_part_n + (size_t)D.nblk0 * 8;
        Ctl& c = E->h_ctl[j];
        std::memset(&c, 0, sizeof(Ctl));
        c.total_time = Jb.model.total_time;
        c.dts_min = std::numeric_limits<double>::infinity();
        c.step_limit = std::numeric_limits<long long>::max();
        c.ke0 = Jb.ke0;
        c.ke = Jb.ke0;
        c.grav_amp = 1.0;
    }
    if ((st = write_ctl(E.get()))) return st;
    if ((st = upload_jobs(E.get()))) return st;
    const bool mixed = E->precision == ENSC_MIXED;
    E->ke = mixed ? pick_e<true>(E->blk_e) : pick_e<false>(E->blk_e);
    E->kn = mixed ? pick_n<true>(E->blk_n) : pick_n<false>(E->blk_n);
    *out = E.release();
    return ENSC_OK;
}

ensc_status ensc_engine_destroy(ensc_engine* E) {
    if (!E) return ENSC_OK;
    (void)hipSetDevice(E->device);
    (void)hipStreamSynchronize(E->stream);
    for (hipEvent_t ev : E->evs) (void)hipEventDestroy(ev);
    if (E->h_ctl) (void)hipHostFree(E->h_ctl);
    E->jobs.clear();
    if (E->stream) (void)hipStreamDestroy(E->stream);
    delete E;
    return ENSC_OK;
}

ensc_status ensc_engine_init(ensc_engine* E) {
    if (!E) return fail(ENSC_ERR_ARG, "null engine");
    HIPCHK(hipSetDevice(E->device));
    ensc_status st;
    if ((st = sync(E)) || (st = read_ctl(E))) return st;
    bool any = false;
    for (size_t j = 0; j < E->jobs.size(); ++j) {
        if (E->jobs[j]->inited) continue;
        E->h_ctl[j].init_pending = 1;
        E->jobs[j]->inited = true;
        any = true;
    }
    if (!any) return ENSC_OK;  // init() runs once (exDyna3D.cpp:109-116)
    if ((st = write_ctl(E))) return st;
    if ((st = enqueue_ticks(E, 1))) return st;
    return sync(E);
}

ensc_status ensc_engine_set_step(ensc_engine* E, uint32_t si) {
    if (!E) return fail(ENSC_ERR_ARG, "null engine");
    HIPCHK(hipSetDevice(E->device));
    ensc_status st;
    if ((st = sync(E))) return st;
    bool changed = false;
    for (auto& Jp : E->jobs) {
        Job& J = *Jp;
        if (!J.inited) return fail(ENSC_ERR_STATE, "ensc_engine_init must run before set_step");
        if (!job_set_step(J, si)) continue;
        changed = true;
        std::vector<uint8_t> mask;
        std::vector<double> val;
        std::vector<int16_t> amp;
        std::vector<uint32_t> nodes;
        job_bc_arrays(J, mask, val, amp, nodes);
        HIPCHK(hipMemcpyAsync(J.d_bcmask, mask.data(), mask.size(), hipMemcpyHostToDevice, E->stream));
        HIPCHK(hipMemcpyAsync(J.d_bcval, val.data(), val.size() * sizeof(double), hipMemcpyHostToDevice, E->stream));
        HIPCHK(hipMemcpyAsync(J.d_bcamp, amp.data(), amp.size() * sizeof(int16_t), hipMemcpyHostToDevice, E->stream));
        HIPCHK(hipMemsetAsync(J.dev.rf, 0, 3 * (size_t)J.model.Nn * sizeof(double), E->stream));
        if (nodes.size() > J.d_bcnodes_cap) {
            J.d_bcnodes = static_cast<uint32_t*>(J.arena.alloc(nodes.size() * 4));
            if (!J.d_bcnodes) return fail(ENSC_ERR_MEM, "BC node list allocation failed");
            J.d_bcnodes_cap = nodes.size();
        }
        if (!nodes.empty())
            HIPCHK(hipMemcpyAsync(J.d_bcnodes, nodes.data(), nodes.size() * 4, hipMemcpyHostToDevice, E->stream));
        const GravState& g = J.grav[si];
        J.dev.grav_on = g.on;
        J.dev.grav_amp = g.amp;
        J.dev.grav_g = g.g;
        for (int d = 0; d < 3; ++d) J.dev.grav_dir[d] = g.dir[d];
        J.dev.bcmask = J.d_bcmask;
    }
    if (!changed) return ENSC_OK;
    if ((st = upload_jobs(E))) return st;
    for (size_t j = 0; j < E->jobs.size(); ++j) {
        Job& J = *E->jobs[j];
        if (J.cur_step != si) continue;
        uint32_t n = 0;
        std::vector<uint8_t> dummy;
        // recount BC nodes (same list as uploaded above)
        std::vector<uint8_t> mask;
        std::vector<double> val;
        std::vector<int16_t> amp;
        std::vector<uint32_t> nodes;
        job_bc_arrays(J, mask, val, amp, nodes);
        n = (uint32_t)nodes.size();
        const uint32_t blocks = std::max<uint32_t>(1, (n + 255) / 256);
        hipLaunchKernelGGL(k_apply_bc, dim3(blocks), dim3(256), 0, E->stream, E->d_jobs, (uint32_t)j,
                           (const uint32_t*)J.d_bcnodes, n, J.model.steps[si].period, 1);
    }
    HIPCHK(hipGetLastError());
    if ((st = sync(E)) || (st = read_ctl(E))) return st;
    for (size_t j = 0; j < E->jobs.size(); ++j) E->h_ctl[j].step_index = (int32_t)E->jobs[j]->cur_step;
    return write_ctl(E);
}

ensc_status ensc_engine_advance(ensc_engine* E, int64_t max_steps, int32_t* done) {
    if (!E) return fail(ENSC_ERR_ARG, "null engine");
    HIPCHK(hipSetDevice(E->device));
    ensc_status st;
    if ((st = sync(E)) || (st = read_ctl(E))) return st;
    for (size_t j = 0; j < E->jobs.size(); ++j)
        E->h_ctl[j].step_limit = max_steps > 0 ? E->h_ctl[j].nsteps + max_steps : std::numeric_limits<long long>::max();
    if ((st = write_ctl(E))) return st;
    int64_t chunk = 4;
    for (;;) {
        bool all = true;
        for (size_t j = 0; j < E->jobs.size(); ++j) all = all && job_done(E->h_ctl[j]);
        if (all) break;
        const bool t = E->timing;
        E->timing = false;
        st = enqueue_ticks(E, chunk);
        E->timing = t;
        if (st) return st;
        if ((st = read_ctl(E))) return st;
        chunk = std::min<int64_t>(chunk * 2, 256);
    }
    bool step_end = true, errs = false;
    for (size_t j = 0; j < E->jobs.size(); ++j) {
        step_end = step_end && !(E->h_ctl[j].time < E->h_ctl[j].step_end);
        errs = errs || E->h_ctl[j].err != 0;
        E->h_ctl[j].step_limit = std::numeric_limits<long long>::max();
    }
    if ((st = write_ctl(E))) return st;
    if (done) *done = step_end ? 1 : 0;
    if (errs) return fail(ENSC_ERR_NUMERIC, "a job raised a device error flag (see ENSC_S_ERR)");
    return ENSC_OK;
}

ensc_status ensc_engine_run(ensc_engine* E) {
    if (!E) return fail(ENSC_ERR_ARG, "null engine");
    ensc_status st;
    if ((st = ensc_engine_init(E))) return st;
    const size_t nsteps = E->jobs[0]->model.steps.size();
    for (uint32_t s = 1; s < nsteps; ++s) {
        if ((st = ensc_engine_set_step(E, s))) return st;
        int32_t done = 0;
        if ((st = ensc_engine_advance(E, 0, &done))) return st;
    }
    return ENSC_OK;
}

ensc_status ensc_engine_enqueue_steps(ensc_engine* E, int64_t n) {
    if (!E) return fail(ENSC_ERR_ARG, "null engine");
    HIPCHK(hipSetDevice(E->device));
    if (E->pending_ticks) {
        ensc_status st = sync(E);
        if (st) return st;
    }
    return enqueue_ticks(E, n);
}

ensc_status ensc_engine_sync(ensc_engine* E) {
    if (!E) return fail(ENSC_ERR_ARG, "null engine");
    HIPCHK(hipSetDevice(E->device));
    return sync(E);
}

ensc_status ensc_engine_download(ensc_engine* E, const char* field, double* out, int64_t count) {
    if (!E || !field || !out) return fail(ENSC_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(E->device));
    ensc_status st = sync(E);
    if (st) return st;
    Job& J = *E->jobs[0];
    const size_t Nn = J.model.Nn, Ne = J.model.Ne, es = J.ne_s;
    const bool mixed = E->precision == ENSC_MIXED;
    const std::string f(field);
    if (f == "u" || f == "v" || f == "x") {
        if (count < (int64_t)(3 * Nn)) return fail(ENSC_ERR_ARG, "buffer too small");
        std::vector<double> rec(Nn * NODE_REC);
        HIPCHK(hipMemcpy(rec.data(), J.dev.node, rec.size() * sizeof(double), hipMemcpyDeviceToHost));
        const int off = f == "x" ? 0 : (f == "v" ? 3 : 6);
        for (size_t i = 0; i < Nn; ++i) {
            const size_t o = J.n_new2old[i];
            for (int d = 0; d < 3; ++d) out[3 * o + d] = rec[i * NODE_REC + off + d];
        }
        return ENSC_OK;
    }
    if (f == "rf") {
        if (count < (int64_t)(3 * Nn)) return fail(ENSC_ERR_ARG, "buffer too small");
        std::vector<double> r(3 * Nn);
        HIPCHK(hipMemcpy(r.data(), J.dev.rf, r.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < Nn; ++i)
            for (int d = 0; d < 3; ++d) out[3 * (size_t)J.n_new2old[i] + d] = r[3 * i + d];
        return ENSC_OK;
    }
    int nc = 0;
    const void* src = nullptr;
    if (f == "sigma") {
        nc = 6;
        src = J.dev.sigma;
    } else if (f == "qia") {
        nc = 12;
        src = J.dev.qia;
    } else if (f == "epsp") {
        nc = 1;
        src = J.dev.epsp;
    } else {
        return fail(ENSC_ERR_ARG, "unknown field '" + f + "'");
    }
    if (count < (int64_t)(nc * Ne)) return fail(ENSC_ERR_ARG, "buffer too small");
    std::vector<double> tmp((size_t)nc * es);
    if (mixed) {
        std::vector<float> t32((size_t)nc * es);
        HIPCHK(hipMemcpy(t32.data(), src, t32.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < t32.size(); ++i) tmp[i] = t32[i];
    } else {
        HIPCHK(hipMemcpy(tmp.data(), src, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
    for (size_t i = 0; i < Ne; ++i) {
        const size_t o = J.e_new2old[i];
        for (int q = 0; q < nc; ++q) out[(size_t)nc * o + q] = tmp[(size_t)q * es + i];
    }
    return ENSC_OK;
}

ensc_status ensc_engine_scalars(ensc_engine* E, double* out) {
    if (!E || !out) return fail(ENSC_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(E->device));
    ensc_status st;
    if ((st = sync(E)) || (st = read_ctl(E))) return st;
    const Ctl& c = E->h_ctl[0];
    out[ENSC_S_TIME] = c.time;
    out[ENSC_S_DT_I] = c.dt_i;
    out[ENSC_S_DT_I_1] = c.dt_i_1;
    out[ENSC_S_DTS_MIN] = c.dts_min;
    out[ENSC_S_KE] = c.ke;
    out[ENSC_S_IE] = c.ie;
    out[ENSC_S_HG] = c.hg;
    out[ENSC_S_VIS] = c.vis;
    out[ENSC_S_WEXT] = c.wext;
    out[ENSC_S_KE0] = c.ke0;
    out[ENSC_S_NSTEPS] = (double)c.nsteps;
    out[ENSC_S_STEP_INDEX] = (double)c.step_index;
    out[ENSC_S_STEP_END] = c.step_end;
    out[ENSC_S_ERR] = (double)c.err;
    out[ENSC_S_DONE] = (c.time < c.step_end) ? 0.0 : 1.0;
    return ENSC_OK;
}

ensc_status ensc_engine_history(ensc_engine* E, double* out, int64_t rows, int64_t* got) {
    if (!E || !got) return fail(ENSC_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(E->device));
    ensc_status st;
    if ((st = sync(E)) || (st = read_ctl(E))) return st;
    Job& J = *E->jobs[0];
    int64_t n = std::min<int64_t>(E->h_ctl[0].nsteps, J.hist_cap);
    n = std::min<int64_t>(n, rows);
    *got = n;
    if (n > 0 && out)
        HIPCHK(hipMemcpy(out, J.d_hist, (size_t)n * (ENSC_H_BASE + 3 * J.dev.n_probe) * sizeof(double),
                         hipMemcpyDeviceToHost));
    return ENSC_OK;
}

ensc_status ensc_engine_kernel_times(ensc_engine* E, double* elem_ms, double* node_ms, int64_t* launches) {
    if (!E) return fail(ENSC_ERR_ARG, "null engine");
    ensc_status st = sync(E);
    if (st) return st;
    if (elem_ms) *elem_ms = E->elem_ms;
    if (node_ms) *node_ms = E->node_ms;
    if (launches) *launches = E->launches;
    return ENSC_OK;
}

ensc_status ensc_engine_set_timing(ensc_engine* E, int32_t on) {
    if (!E) return fail(ENSC_ERR_ARG, "null engine");
    ensc_status st = sync(E);
    if (st) return st;
    E->timing = on != 0;
    return ENSC_OK;
}

int64_t ensc_engine_device_bytes(ensc_engine* E) {
    if (!E) return 0;
    int64_t t = (int64_t)E->arena.total;
    for (auto& j : E->jobs) t += (int64_t)j->arena.total;
    return t;
}

}  // extern "C"
