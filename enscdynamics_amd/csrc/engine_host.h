// Host-side structures of the step engine.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>

#include "guard.h"
#include <string>
#include <vector>

#include "../../include/ensc_engine.h"
#include "engine_dev.h"

namespace ensc {

extern thread_local std::string g_last_error;
ensc_status fail(ensc_status s, const std::string& msg);
size_t round_up(size_t v, size_t m);

struct BcCopy {
    std::vector<uint32_t> nodes;  // original numbering
    int32_t d0 = 0, d1 = 0, cond = 1, amp = -1;
    double value = 0.0;
};

struct DsloadCopy {
    std::vector<uint32_t> elems;  // original numbering
    std::vector<int32_t> faces;   // 0..5 = S1..S6
    int32_t amp = -1;
    double value = 0.0;
};

struct StepCopy {
    double period = 0.0;
    int32_t reset_spc = 0, reset_vel = 0, reset_dload = 0, grav_on = 0, grav_amp = -1, reset_dsload = 0;
    double grav_value = 0.0, grav_dir[3] = {0, 0, 0};
    std::vector<BcCopy> spc, vel;
    std::vector<DsloadCopy> dsload;
};

// The caller's model, copied (include/ensc_model.h).
struct ModelCopy {
    uint32_t Nn = 0, Ne = 0;
    std::vector<double> x0, v_init;
    std::vector<uint32_t> conn;
    double E = 0, nu = 0, rho = 0, G = 0, K = 0, lambda = 0, wos = 0;
    int32_t plastic = 0;
    double yield0 = 0, hard = 0, c_dr = 0, c_cr = 0, factor = 0, cvisl = 0, cvisq = 0, total_time = 0;
    std::vector<StepCopy> steps;
    std::vector<double> amp_t, amp_a;
    std::vector<int32_t> amp_off;
};

struct BcRef {
    uint32_t step, idx;
};

struct GravState {
    int32_t on = 0, amp = -1;
    double g = 0.0, dir[3] = {0, 0, 0};
};

struct DeviceArena {
    std::vector<void*> ptrs;
    size_t total = 0;
    void* alloc(size_t bytes);
    ~DeviceArena();
};

// Host-side tables of a job, ready to upload (prepare_job -> upload_job).
struct JobImage {
    int precision = 0, blk_e = 256, blk_n = 256;
    uint32_t W = 0;
    size_t es = 0, ns = 0;
    std::vector<uint32_t> conn_d, ell, blk_ptr, blk_nodes, pos_ptr, pos_list;
    std::vector<uint16_t> lconn;
    std::vector<double> gi64, v064, pos, vu;  // gi64, v064: FP64 engine
};

struct Job {
    ModelCopy model;
    JobDev dev{};
    DeviceArena arena;
    std::vector<uint32_t> e_new2old, e_old2new, n_new2old, n_old2new;
    size_t ne_s = 0, nn_s = 0;
    uint32_t ell_w = 0;
    uint32_t lds_nodes = 0;  // largest distinct-node count of an element block
    bool asm_capable = false; // every element block free of corner conflicts (block assembly possible)
    int precision = ENSC_FP64;  // the job's own precision (job_precision)
    std::vector<uint32_t> blk_ptr_h;  // host copy of the staging list offsets (block maps)
    std::unique_ptr<JobImage> img;    // between prepare_job and upload_job
    double ke0 = 0.0;
    // setCurrentStep state (exDyna3D.h:215-220)
    uint32_t cur_step = 0;
    std::vector<BcRef> cur_spc, cur_vel, prev_spc, prev_vel;
    std::vector<GravState> grav;
    int32_t dsl_step = -1;        // step whose *Dsload lines are current (-1: none)
    uint8_t* d_bcmask = nullptr;
    double* d_bcval = nullptr;
    int16_t* d_bcamp = nullptr;
    uint32_t* d_bcnodes = nullptr;
    size_t d_bcnodes_cap = 0;
    std::vector<uint32_t> probes;
    double* d_hist = nullptr;
    long long hist_cap = 0;
    bool inited = false;
    bool pending_step = false;    // started by the serving loop, step 1 not switched in yet
    bool keep_in_layout = false;  // single-job engines keep their job in every launch
    JobDev* d_self = nullptr;     // device copy of this job's table entry (BC kernel)
    uint32_t snap_slots = 0;      // output snapshot ring allocated (slots)
    double* d_post = nullptr;     // [ne][13] output fields of one snapshot (k_post)
    JobDev self_copy{};
};

// Virtual particles of the FSI boundary (host side of the coupling): the
// elements that received particles, ascending (original numbering), the
// particle range of each and every particle's unit coordinates.
struct FsiParticles {
    std::vector<uint32_t> elems, eptr;
    std::vector<double> unit;  // [n][3]
};
void fsi_particles(const ModelCopy& M, const std::vector<double>& x, int n_layers, double dist, FsiParticles& out);

bool copy_model(const ensc_model_desc* m, ModelCopy& c, std::string& err);
ensc_status prepare_job(Job& J, int precision, int reorder, int blk_e, int blk_n, int fp64_rule);
ensc_status upload_job(Job& J);
// The precision a job runs in, from the engine's precision and the rule of
// ensc_options.fp64_jobs (0 auto, 1 mixed, 2 FP64), reading the job alone.
int job_precision(const ModelCopy& m, int engine_precision, int fp64_rule);
constexpr double QUASI_STATIC_TRANSITS = 512.0;
bool job_set_step(Job& J, uint32_t si);
void job_bc_arrays(const Job& J, std::vector<uint8_t>& mask, std::vector<double>& val, std::vector<int16_t>& amp,
                   std::vector<uint32_t>& nodes);
void job_dsload_arrays(const Job& J, std::vector<uint32_t>& fnodes, std::vector<double>& fval,
                       std::vector<int32_t>& famp, std::vector<uint32_t>& nptr, std::vector<uint32_t>& nlist);

}  // namespace ensc
