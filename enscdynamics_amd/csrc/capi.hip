// C ABI of the step engine (include/ensc_engine.h).
//
// An engine holds jobs in stable slots.  Every live job is laid out in one
// job table with contiguous element-block and node-block ranges, so one
// launch of each kernel advances all of them ("tick").  The layout is
// rebuilt between ticks whenever jobs are added, retired or finish.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "engine_host.h"
#include "../../include/ensc_vtu.h"
#include "kernels.h"

using namespace ensc;

#define HIPCHK(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return fail(ENSC_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

typedef void (*KernelFn)(const JobDev*, const uint4*);

namespace {

// Size of the static scratch array of ensc_engine_advance_static: the job
// table (~0.6 KB per job), the block maps (16 B per block) and the block
// partials (40 B per element block, 16 B per node block) of one engine; 4 MiB
// hold a single mesh of ~10M elements.
constexpr size_t STATIC_SCRATCH_BYTES = (size_t)4 << 20;
// Element finalize: one block for a job of up to FIN_ROLE_MIN block partials
// (element or node); up to FIN_ONE, four blocks that each fold the partials
// of the scalars only they write (FIN_ROLES); else one block per FIN_SLICE
// partials, at most FIN_MAX_G, each folding a slice of every value, the last
// to count deciding; per job FIN_JOB_BYTES of device memory hold the slices'
// folds and the ticket.  Measured on the 1M-element J2 cube: the ticket
// (fences and atomic) costs ~1 us more than the split fold saves at 4k
// partials, and saves ~3 us at 7.8k (8 blocks); the four role blocks save
// ~2.5 us at 4k.  Small jobs keep one block: a 512-thread finalize block
// takes a whole CU (160 VGPRs), so four per job in a batch of 256 meshes of
// 1k-50k elements ran in four rounds (7.8 -> 11 us per tick).
constexpr uint32_t FIN_ROLE_MIN = 1024, FIN_ONE = 4096, FIN_SLICE = 1024, FIN_MAX_G = 8, FIN_ROLES = 4;
constexpr size_t FIN_JOB_BYTES = FIN_MAX_G * 8 * sizeof(double) + 256;

// Device buffer that grows (contents are not preserved).
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    bool ensure(size_t bytes) {
        if (bytes <= cap && p) return true;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        const size_t b = std::max<size_t>(round_up(bytes, 4096), 4096);
        if (hipMalloc(&p, b) != hipSuccess) return false;
        cap = b;
        return true;
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

struct ensc_engine {
    int device = 0, precision = ENSC_FP64, blk_e = 256, blk_n = 256, reorder = 0, staging_opt = 0, waves = 2;
    int fp64_rule = 0;  // ensc_options.fp64_jobs (per-job precision of a mixed engine)
    int fp64_order = 0; // ensc_options.fp64_order (0: block partials, 1: the reference's order)
    int waves64 = 1;    // occupancy target of the FP64 staged element kernel (reference-order body)
    int waves_w64 = 2;  // occupancy target of the FP64 block-assembly element kernel
    bool pipe_opt = true;  // staging option 0: pipelined element pass where it applies
    hipStream_t stream = nullptr;
    // second stream for the staged element blocks of a tick that also runs
    // the pipelined pass (forked from and joined to `stream` by events)
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::vector<std::unique_ptr<Job>> slots;  // job id = slot index; null = free
    // per-slot scalars (device array + pinned host mirror), never moved while jobs live
    Ctl* d_ctl = nullptr;
    Ctl* h_ctl = nullptr;
    uint32_t ctl_cap = 0;
    // layout of the jobs currently stepped (relayout): the element block
    // map holds, in this order, the blocks of the pipelined pass, then the
    // staged and the gathered blocks of mixed-precision jobs, then the staged
    // and the gathered blocks of FP64 jobs; the node block map the mixed
    // block-assembly jobs' blocks, the other mixed jobs', the FP64
    // block-assembly jobs' and the other FP64 jobs'
    std::vector<uint32_t> live;
    DevBuf jobs_buf, b2e_buf, b2n_buf, b2f_buf, pe_buf, pn_buf, b2fin_buf, fin_buf;
    // where the kernels find the job table and block maps: the DevBufs, or
    // the static scratch array (ensc_engine_advance_static)
    void *jobs_p = nullptr, *b2e_p = nullptr, *b2n_p = nullptr, *b2f_p = nullptr, *b2fin_p = nullptr;
    uint32_t nblk_f = 0;    // face blocks of the surface-pressure kernel
    uint32_t nblk_fin = 0;  // blocks of the element finalize (FIN_ROLES or fin_g per job)
    uint32_t fsi_np = 0, fsi_nn = 0;  // virtual particles and FSI nodes of the single job
    bool in_scratch = false;      // the call in progress places its tables in scratch
    bool layout_scratch = false;  // the current layout lives in scratch
    char* scratch_base = nullptr;
    uint32_t nblk_pipe = 0, nblk_e = 0, nblk_g = 0;  // element blocks: pipelined, staged, gathered (mixed)
    bool pipe_bc = false;  // the pipelined launch holds tiles of a job with amplitude velocity BCs
    uint32_t nblk_e64 = 0, nblk_g64 = 0;             // element blocks: staged, gathered (FP64)
    uint32_t nblk_w64 = 0;  // element blocks of FP64 block-assembly jobs (Walsh-form body)
    uint32_t nblk_n_asm = 0, nblk_n = 0;      // node blocks, mixed jobs: block partials, per-corner forces
    uint32_t nblk_n64_asm = 0, nblk_n64 = 0;  // node blocks, FP64 jobs: block partials, per-corner forces
    bool dirty = true;
    KernelFn ke = nullptr, ke_g = nullptr;      // k_element staged / gathered, mixed jobs
    KernelFn ke64 = nullptr, ke64_g = nullptr;  // the same, FP64 jobs
    KernelFn ke64w = nullptr;                   // FP64 block-assembly jobs
    KernelFn kn = nullptr, kn_asm = nullptr;      // k_node, mixed jobs: per-corner / block partials
    KernelFn kn64 = nullptr, kn64_asm = nullptr;  // k_node, FP64 jobs
    bool staged = false;      // some live job stages its nodes in LDS
    uint32_t n_asm = 0;       // live jobs using block assembly
    uint32_t n_fp64 = 0;      // live jobs running in FP64
    uint32_t pipe_grid = 0;   // the pipelined kernel's block count (pipe_per_cu per CU)
    int ncu = 0, pipe_per_cu = 2;
    size_t lds_bytes = 0, lds_bytes64 = 0, lds_bytes_w64 = 0;  // dynamic LDS of the staged k_element (mixed, FP64, FP64 block assembly)
    // output snapshots of a single-job engine: files go to out_dir (empty:
    // the caller drains them), numbered from out_count
    std::string out_dir;
    int64_t out_count = 0;
    // timing
    bool timing = false;
    std::vector<hipEvent_t> evs;
    double elem_ms = 0.0, node_ms = 0.0, phase_ms[4] = {0.0, 0.0, 0.0, 0.0};
    int64_t launches = 0, pending_ticks = 0;
    // graph of graph_ticks ticks, valid for the current layout
    int32_t graph_ticks = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    bool graph_valid = false;
};

namespace {

// Plain device copy, 16 bytes per lane per iteration: the reference kernel
// for HBM bandwidth and for scaling hardware byte counters.
__global__ __launch_bounds__(256) void k_copy_probe(const double2* __restrict__ a, double2* __restrict__ b, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) b[i] = a[i];
}

// Zero fill of a small device array (the finalize tickets of a new layout);
// a kernel, not a memset node, so it is safe wherever it is captured.
__global__ __launch_bounds__(256) void k_zero(double* __restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = 0.0;
}

template <bool M, int W>
KernelFn pick_staged(int blk) {
    switch (blk) {
        case 64: return k_element<M, 64, true, W>;
        case 128: return k_element<M, 128, true, W>;
        default: return k_element<M, 256, true, W>;
    }
}
// Element kernel for (precision, block size, staging, occupancy target).
// The occupancy target (waves per SIMD) caps the kernel's registers; it is
// a tuning choice only and never changes results.
template <bool M>
KernelFn pick_e(int blk, bool staged, int waves) {
    if (!staged) {
        switch (blk) {
            case 64: return k_element<M, 64, false, 1>;
            case 128: return k_element<M, 128, false, 1>;
            default: return k_element<M, 256, false, 1>;
        }
    }
    switch (waves) {
        case 1: return pick_staged<M, 1>(blk);
        case 3: return pick_staged<M, 3>(blk);
        default: return pick_staged<M, 2>(blk);
    }
}
// Element kernel of FP64 block-assembly jobs (the Walsh-form FP64 body).
template <int W>
KernelFn pick_w64_w(int blk) {
    switch (blk) {
        case 64: return k_element<false, 64, true, W, true>;
        case 128: return k_element<false, 128, true, W, true>;
        default: return k_element<false, 256, true, W, true>;
    }
}
KernelFn pick_w64(int blk, int waves) {
    switch (waves) {
        case 1: return pick_w64_w<1>(blk);
        case 3: return pick_w64_w<3>(blk);
        default: return pick_w64_w<2>(blk);
    }
}
// Node kernel for (precision, block size, force source): ASM sums the block
// partials of block-assembly jobs, otherwise the per-corner forces.
template <bool M, bool ASM>
KernelFn pick_n(int blk) {
    switch (blk) {
        case 64: return k_node<M, 64, ASM>;
        case 128: return k_node<M, 128, ASM>;
        default: return k_node<M, 256, ASM>;
    }
}

Job* job_at(ensc_engine* E, uint32_t id) {
    return id < E->slots.size() ? E->slots[id].get() : nullptr;
}

ensc_status read_ctl(ensc_engine* E) {
    if (!E->ctl_cap) return ENSC_OK;
    HIPCHK(hipMemcpyAsync(E->h_ctl, E->d_ctl, E->ctl_cap * sizeof(Ctl), hipMemcpyDeviceToHost, E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
    return ENSC_OK;
}

ensc_status write_ctl(ensc_engine* E) {
    if (!E->ctl_cap) return ENSC_OK;
    HIPCHK(hipMemcpyAsync(E->d_ctl, E->h_ctl, E->ctl_cap * sizeof(Ctl), hipMemcpyHostToDevice, E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
    return ENSC_OK;
}

bool job_done(const Ctl& c) { return !(c.time < c.step_end) || c.nsteps >= c.step_limit || c.err != 0; }

// Grow the slot scalar array to hold `need` slots (device idle).
ensc_status grow_ctl(ensc_engine* E, uint32_t need) {
    if (need <= E->ctl_cap) return ENSC_OK;
    const uint32_t cap = std::max<uint32_t>(need, std::max<uint32_t>(16, 2 * E->ctl_cap));
    Ctl* d = nullptr;
    Ctl* h = nullptr;
    HIPCHK(hipMalloc((void**)&d, cap * sizeof(Ctl)));
    HIPCHK(hipHostMalloc((void**)&h, cap * sizeof(Ctl), hipHostMallocDefault));
    std::memset(h, 0, cap * sizeof(Ctl));
    if (E->ctl_cap) {
        ensc_status st = read_ctl(E);
        if (st) return st;
        std::memcpy(h, E->h_ctl, E->ctl_cap * sizeof(Ctl));
        (void)hipFree(E->d_ctl);
        (void)hipHostFree(E->h_ctl);
    }
    E->d_ctl = d;
    E->h_ctl = h;
    E->ctl_cap = cap;
    for (size_t s = 0; s < E->slots.size(); ++s)
        if (E->slots[s]) E->slots[s]->dev.ctl = E->d_ctl + s;
    E->dirty = true;
    return write_ctl(E);
}

// Launched block h runs on XCD h % NXCD (the dispatcher deals workgroups to
// the XCDs round robin).  Give each XCD a contiguous range of the logical
// blocks, so blocks that share nodes (neighbours in the space-filling order)
// meet in the same L2.  A permutation: results do not depend on it.
constexpr uint32_t NXCD = 8;
std::vector<uint4> xcd_order(const std::vector<uint4>& logical) {
    const size_t n = logical.size(), q = n / NXCD, r = n % NXCD;
    std::vector<uint4> out(n);
    for (size_t h = 0; h < n; ++h) {
        const size_t x = h % NXCD, i = h / NXCD;
        out[h] = logical[x * q + std::min(x, r) + i];
    }
    return out;
}

// Rebuild the job table and block maps for the jobs that are not finished.
ensc_status relayout(ensc_engine* E) {
    ensc_status st;
    if ((st = read_ctl(E))) return st;
    E->live.clear();
    for (uint32_t s = 0; s < E->slots.size(); ++s) {
        Job* J = E->slots[s].get();
        if (!J) continue;
        const Ctl& c = E->h_ctl[s];
        const bool idle = J->inited && !c.init_pending && !(c.time < c.step_end) && !J->pending_step;
        if (!idle || J->keep_in_layout) E->live.push_back(s);
    }
    // Every choice below depends on the job alone (and the engine options),
    // so a job runs the same kernels, with the same arithmetic, alone and in
    // any mix:
    //  - precision: the job's own (job_precision);
    //  - staged: its largest block node list fits 64 KiB of LDS (72 B per
    //    node, plus the block assembly accumulators: 12 B mixed, 24 B FP64);
    //  - block assembly: staged and free of corner conflicts (FP64 jobs:
    //    unless the reference's summation order is asked for);
    //  - pipelined pass: block assembly and tiles of PIPE_BLK elements; each
    //    such tile with fewer than PIPE_CAP nodes (for a job with amplitude
    //    velocity BCs, the pass overrides the staged velocities of the tiles
    //    that hold a constrained node)
    //    (the rest of the job's tiles use the staged pass).  The split
    //    follows from the job's own tiles, so it too is the same in any mix.
    auto asm_wanted = [&](const Job& J) { return J.precision == ENSC_MIXED || E->fp64_order == 0; };
    auto lds_per_node = [&](const Job& J) {
        return 9 * sizeof(double) + (asm_wanted(J) ? (J.precision == ENSC_MIXED ? 12 : 24) : 0);
    };
    auto job_staged = [&](const Job& J) {
        return E->staging_opt != 2 && (size_t)J.lds_nodes * lds_per_node(J) <= 65536;
    };
    auto job_asm = [&](const Job& J) { return asm_wanted(J) && J.asm_capable && job_staged(J); };
    auto job_pipe = [&](const Job& J) {
        return E->pipe_opt && job_asm(J) && E->blk_e == PIPE_BLK;
    };
    // block maps: (job, block of the job, staged range) per launched block,
    // in XCD order, grouped as described at ensc_engine::live
    std::vector<uint4> lp, ls, lg, ls64, lw64, lg64, lna, ln, lna64, ln64, lf, lfin;
    std::vector<JobDev> tab;
    std::vector<size_t> pe0, pn0;
    size_t ne_blocks = 0, n_blocks = 0;
    uint32_t lds_nodes = 0, lds_nodes64 = 0, lds_nodes_w64 = 0;
    bool any_asm = false, any_asm64 = false;
    E->n_asm = 0;
    E->n_fp64 = 0;
    for (uint32_t i = 0; i < E->live.size(); ++i) {
        Job& J = *E->slots[E->live[i]];
        const JobDev& D = J.dev;
        pe0.push_back(ne_blocks);
        pn0.push_back(n_blocks);
        const bool m = J.precision == ENSC_MIXED;
        const bool as = job_asm(J), pipe = job_pipe(J), staged = job_staged(J);
        for (uint32_t b = 0; b < D.nblk_e; ++b) {
            const uint32_t nodes = J.blk_ptr_h[b + 1] - J.blk_ptr_h[b];
            std::vector<uint4>* le;
            if (!m)
                le = as ? &lw64 : (staged ? &ls64 : &lg64);
            else
                le = (pipe && nodes < (uint32_t)PIPE_CAP) ? &lp : (staged ? &ls : &lg);
            if (le == &ls) {
                lds_nodes = std::max(lds_nodes, nodes);
                any_asm |= as;
            } else if (le == &ls64) {
                lds_nodes64 = std::max(lds_nodes64, nodes);
                any_asm64 |= as;
            } else if (le == &lw64) {
                lds_nodes_w64 = std::max(lds_nodes_w64, nodes);
            }
            le->push_back(make_uint4(i, b, J.blk_ptr_h[b], nodes));
        }
        ne_blocks += D.nblk_e;
        std::vector<uint4>& lnj = !m ? (as ? lna64 : ln64) : (as ? lna : ln);
        for (uint32_t b = 0; b < D.nblk_n; ++b) lnj.push_back(make_uint4(i, b, 0, 0));
        n_blocks += D.nblk_n;
        for (uint32_t b = 0; b < (D.n_dsf + FACE_BLK - 1) / FACE_BLK; ++b) lf.push_back(make_uint4(i, b, 0, 0));
        // element finalize blocks (the job alone decides, so its sums' order
        // is the same in any mix)
        const uint32_t nmax = std::max(D.nblk_e, D.nblk_n);
        const uint32_t g =
            nmax <= FIN_ONE ? 1u : std::min<uint32_t>(FIN_MAX_G, (nmax + FIN_SLICE - 1) / FIN_SLICE);
        J.dev.fin_g = g;
        if (g == 1 && nmax > FIN_ROLE_MIN)
            for (uint32_t r = 1; r <= FIN_ROLES; ++r) lfin.push_back(make_uint4(i, 0, 1, r));  // FIN_DECIDE .. FIN_WORK_B
        else
            for (uint32_t q = 0; q < g; ++q) lfin.push_back(make_uint4(i, q, g, 0));  // FIN_ALL
        J.dev.block_asm = as ? 1 : 0;
        E->n_asm += as;
        E->n_fp64 += !m;
    }
    std::vector<uint4> b2e, b2n;
    for (const std::vector<uint4>* l : {&lp, &ls, &lg, &ls64, &lw64, &lg64}) {
        const std::vector<uint4> r = xcd_order(*l);
        b2e.insert(b2e.end(), r.begin(), r.end());
    }
    for (const std::vector<uint4>* l : {&lna, &ln, &lna64, &ln64}) {
        const std::vector<uint4> r = xcd_order(*l);
        b2n.insert(b2n.end(), r.begin(), r.end());
    }
    E->nblk_pipe = (uint32_t)lp.size();
    E->pipe_bc = false;
    for (const uint4& q : lp) E->pipe_bc |= E->slots[E->live[q.x]]->dev.dyn_bc != 0;
    E->nblk_e = (uint32_t)ls.size();
    E->nblk_g = (uint32_t)lg.size();
    E->nblk_e64 = (uint32_t)ls64.size();
    E->nblk_w64 = (uint32_t)lw64.size();
    E->nblk_g64 = (uint32_t)lg64.size();
    E->nblk_n_asm = (uint32_t)lna.size();
    E->nblk_n = (uint32_t)ln.size();
    E->nblk_n64_asm = (uint32_t)lna64.size();
    E->nblk_n64 = (uint32_t)ln64.size();
    E->nblk_f = (uint32_t)lf.size();
    E->nblk_fin = (uint32_t)lfin.size();
    const size_t sz[8] = {std::max<size_t>(1, E->live.size()) * sizeof(JobDev),
                          std::max<size_t>(1, b2e.size()) * sizeof(uint4), std::max<size_t>(1, b2n.size()) * sizeof(uint4),
                          std::max<size_t>(1, ne_blocks) * 5 * sizeof(double),
                          std::max<size_t>(1, n_blocks) * 2 * sizeof(double), std::max<size_t>(1, lf.size()) * sizeof(uint4),
                          std::max<size_t>(1, lfin.size()) * sizeof(uint4),
                          std::max<size_t>(1, E->live.size()) * FIN_JOB_BYTES};
    void* pe = nullptr;
    void* pn = nullptr;
    void* fin = nullptr;
    if (E->in_scratch) {
        // the static scratch array: job table, maps and block partials back
        // to back, each 256-byte aligned; nothing is allocated
        size_t need = 0;
        for (size_t b : sz) need += round_up(b, 256);
        if (need > STATIC_SCRATCH_BYTES)
            return fail(ENSC_ERR_ARG, "ensc_engine_advance_static: the launch tables of this engine need " +
                                          std::to_string(need) + " bytes, more than the " +
                                          std::to_string(STATIC_SCRATCH_BYTES) +
                                          "-byte static scratch (use ensc_engine_advance)");
        char* q = E->scratch_base;
        void* part[8];
        for (int k = 0; k < 8; ++k) {
            part[k] = q;
            q += round_up(sz[k], 256);
        }
        E->jobs_p = part[0];
        E->b2e_p = part[1];
        E->b2n_p = part[2];
        pe = part[3];
        pn = part[4];
        E->b2f_p = part[5];
        E->b2fin_p = part[6];
        fin = part[7];
    } else {
        if (!E->jobs_buf.ensure(sz[0]) || !E->b2e_buf.ensure(sz[1]) || !E->b2n_buf.ensure(sz[2]) ||
            !E->pe_buf.ensure(sz[3]) || !E->pn_buf.ensure(sz[4]) || !E->b2f_buf.ensure(sz[5]) ||
            !E->b2fin_buf.ensure(sz[6]) || !E->fin_buf.ensure(sz[7]))
            return fail(ENSC_ERR_MEM, "layout allocation failed");
        E->jobs_p = E->jobs_buf.p;
        E->b2e_p = E->b2e_buf.p;
        E->b2n_p = E->b2n_buf.p;
        E->b2f_p = E->b2f_buf.p;
        E->b2fin_p = E->b2fin_buf.p;
        fin = E->fin_buf.p;
        pe = E->pe_buf.p;
        pn = E->pn_buf.p;
    }
    E->layout_scratch = E->in_scratch;
    for (uint32_t i = 0; i < E->live.size(); ++i) {
        Job& J = *E->slots[E->live[i]];
        J.dev.ctl = E->d_ctl + E->live[i];
        J.dev.part_e = static_cast<double*>(pe) + pe0[i] * 5;
        J.dev.part_n = static_cast<double*>(pn) + pn0[i] * 2;
        char* fj = static_cast<char*>(fin) + (size_t)i * FIN_JOB_BYTES;
        J.dev.fin_part = reinterpret_cast<double*>(fj);
        J.dev.fin_ticket = reinterpret_cast<uint32_t*>(fj + FIN_MAX_G * 8 * sizeof(double));
        tab.push_back(J.dev);
    }
    if (!tab.empty())
        HIPCHK(hipMemcpyAsync(E->jobs_p, tab.data(), tab.size() * sizeof(JobDev), hipMemcpyHostToDevice, E->stream));
    if (!b2e.empty())
        HIPCHK(hipMemcpyAsync(E->b2e_p, b2e.data(), b2e.size() * sizeof(uint4), hipMemcpyHostToDevice, E->stream));
    if (!b2n.empty())
        HIPCHK(hipMemcpyAsync(E->b2n_p, b2n.data(), b2n.size() * sizeof(uint4), hipMemcpyHostToDevice, E->stream));
    if (!lf.empty())
        HIPCHK(hipMemcpyAsync(E->b2f_p, lf.data(), lf.size() * sizeof(uint4), hipMemcpyHostToDevice, E->stream));
    if (!lfin.empty())
        HIPCHK(hipMemcpyAsync(E->b2fin_p, lfin.data(), lfin.size() * sizeof(uint4), hipMemcpyHostToDevice, E->stream));
    // the finalize tickets start at zero (a kernel, not a memset node)
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(256), 0, E->stream, static_cast<double*>(fin),
                       sz[7] / sizeof(double));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(E->stream));
    E->staged = E->nblk_pipe + E->nblk_e + E->nblk_e64 + E->nblk_w64 > 0;
    E->lds_bytes = (size_t)std::max<uint32_t>(lds_nodes, 1) * (any_asm ? 9 * sizeof(double) + 12 : 9 * sizeof(double));
    E->lds_bytes64 = (size_t)std::max<uint32_t>(lds_nodes64, 1) * (9 * sizeof(double) + (any_asm64 ? 24 : 0));
    E->lds_bytes_w64 = (size_t)std::max<uint32_t>(lds_nodes_w64, 1) * (9 * sizeof(double) + 24);
    E->ke = pick_e<true>(E->blk_e, true, E->waves);
    E->ke_g = pick_e<true>(E->blk_e, false, E->waves);
    E->ke64 = pick_e<false>(E->blk_e, true, E->waves64);
    E->ke64_g = pick_e<false>(E->blk_e, false, E->waves64);
    E->ke64w = pick_w64(E->blk_e, E->waves_w64);
    E->kn = pick_n<true, false>(E->blk_n);
    E->kn_asm = pick_n<true, true>(E->blk_n);
    E->kn64 = pick_n<false, false>(E->blk_n);
    E->kn64_asm = pick_n<false, true>(E->blk_n);
    E->dirty = false;
    E->graph_valid = false;
    return ENSC_OK;
}

// One tick = element pass, per-job element finalize (which also closes the
// previous tick's nodal update), nodal update.  A run of ticks ends with one
// per-job node finalize.  With timing on, every tick launches all four
// phases so each is timed on its own.
constexpr int FIN_BLK = 512;
// Returns the first error of the phase's runtime calls (kernel launches
// report theirs through hipGetLastError).
hipError_t launch_phase(ensc_engine* E, int phase) {
    const JobDev* jobs = static_cast<const JobDev*>(E->jobs_p);
    const uint32_t nj = (uint32_t)E->live.size();
    if (!nj) return hipSuccess;
    hipError_t err = hipSuccess;
    auto chk = [&](hipError_t e) {
        if (err == hipSuccess && e != hipSuccess) err = e;
    };
    switch (phase) {
        case 0: {
            // The pipelined pass takes every CU's LDS, so staged and gathered
            // blocks launched beside it (on the side stream) start as its
            // blocks finish and fill its last round instead of following it.
            // The passes cover disjoint element blocks of the tick.
            const bool fork = E->nblk_pipe && (E->nblk_e || E->nblk_g || E->nblk_e64 || E->nblk_w64 || E->nblk_g64);
            hipStream_t s2 = E->stream;
            if (fork) {
                chk(hipEventRecord(E->ev_fork, E->stream));
                chk(hipStreamWaitEvent(E->side, E->ev_fork, 0));
                s2 = E->side;
            }
            const uint4* map = static_cast<const uint4*>(E->b2e_p);
            if (E->nblk_pipe)
                hipLaunchKernelGGL((E->pipe_bc ? k_element_pipe<2, true> : k_element_pipe<2, false>),
                                   dim3(std::min(E->nblk_pipe, E->pipe_grid)), dim3(PIPE_BLK), 0, E->stream, jobs, map,
                                   E->nblk_pipe);
            map += E->nblk_pipe;
            if (E->nblk_e)
                hipLaunchKernelGGL(E->ke, dim3(E->nblk_e), dim3(E->blk_e), E->lds_bytes, s2, jobs, map);
            map += E->nblk_e;
            if (E->nblk_g) hipLaunchKernelGGL(E->ke_g, dim3(E->nblk_g), dim3(E->blk_e), 0, s2, jobs, map);
            map += E->nblk_g;
            if (E->nblk_e64)
                hipLaunchKernelGGL(E->ke64, dim3(E->nblk_e64), dim3(E->blk_e), E->lds_bytes64, s2, jobs, map);
            map += E->nblk_e64;
            if (E->nblk_w64)
                hipLaunchKernelGGL(E->ke64w, dim3(E->nblk_w64), dim3(E->blk_e), E->lds_bytes_w64, s2, jobs, map);
            map += E->nblk_w64;
            if (E->nblk_g64) hipLaunchKernelGGL(E->ke64_g, dim3(E->nblk_g64), dim3(E->blk_e), 0, s2, jobs, map);
            if (fork) {
                chk(hipEventRecord(E->ev_join, E->side));
                chk(hipStreamWaitEvent(E->stream, E->ev_join, 0));
            }
            break;
        }
        case 1:
            hipLaunchKernelGGL(k_fin_element<FIN_BLK>, dim3(E->nblk_fin), dim3(FIN_BLK), 0, E->stream, jobs,
                               static_cast<const uint4*>(E->b2fin_p));
            break;
        case 2: {
            // surface-pressure forces and the virtual particles' nodal forces
            // of the step, then the nodal update, then the particles follow
            // the mesh
            if (E->nblk_f)
                hipLaunchKernelGGL(k_face, dim3(E->nblk_f), dim3(FACE_BLK), 0, E->stream, jobs,
                                   static_cast<const uint4*>(E->b2f_p));
            const bool fsi = nj == 1 && E->slots.size() == 1;
            if (fsi && E->fsi_nn)
                hipLaunchKernelGGL(k_fsi_force, dim3((E->fsi_nn + FSI_BLK - 1) / FSI_BLK), dim3(FSI_BLK), 0, E->stream,
                                   jobs);
            const uint4* map = static_cast<const uint4*>(E->b2n_p);
            if (E->nblk_n_asm) hipLaunchKernelGGL(E->kn_asm, dim3(E->nblk_n_asm), dim3(E->blk_n), 0, E->stream, jobs, map);
            map += E->nblk_n_asm;
            if (E->nblk_n) hipLaunchKernelGGL(E->kn, dim3(E->nblk_n), dim3(E->blk_n), 0, E->stream, jobs, map);
            map += E->nblk_n;
            if (E->nblk_n64_asm)
                hipLaunchKernelGGL(E->kn64_asm, dim3(E->nblk_n64_asm), dim3(E->blk_n), 0, E->stream, jobs, map);
            map += E->nblk_n64_asm;
            if (E->nblk_n64) hipLaunchKernelGGL(E->kn64, dim3(E->nblk_n64), dim3(E->blk_n), 0, E->stream, jobs, map);
            if (fsi && E->fsi_np)
                hipLaunchKernelGGL(k_fsi_particles, dim3((E->fsi_np + FSI_BLK - 1) / FSI_BLK), dim3(FSI_BLK), 0,
                                   E->stream, jobs);
            break;
        }
        default: hipLaunchKernelGGL(k_fin_node<FIN_BLK>, dim3(nj), dim3(FIN_BLK), 0, E->stream, jobs); break;
    }
    return err;
}
hipError_t launch_ticks(ensc_engine* E, int64_t n) {
    hipError_t err = hipSuccess;
    for (int64_t i = 0; i < n && err == hipSuccess; ++i)
        for (int p = 0; p < 3 && err == hipSuccess; ++p) err = launch_phase(E, p);
    if (n > 0 && err == hipSuccess) err = launch_phase(E, 3);
    return err;
}

void drop_graph(ensc_engine* E) {
    if (E->graph_exec) (void)hipGraphExecDestroy(E->graph_exec);
    if (E->graph) (void)hipGraphDestroy(E->graph);
    E->graph_exec = nullptr;
    E->graph = nullptr;
    E->graph_valid = false;
}

// Capture graph_ticks ticks of the current layout.  The previous graph is
// destroyed only once the stream is idle (a replay of it may still be
// queued when set_graph or the autotuner invalidated it), and every call
// made while capturing is checked: a fork/join event call that failed
// would leave the side stream outside the capture, its kernels running at
// capture time instead of in the graph.
ensc_status build_graph(ensc_engine* E) {
    HIPCHK(hipStreamSynchronize(E->stream));
    drop_graph(E);
    HIPCHK(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
    const hipError_t le = launch_ticks(E, E->graph_ticks);
    const hipError_t ke = hipGetLastError();
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(E->stream, &g);
    if (le != hipSuccess || ke != hipSuccess || ee != hipSuccess) {
        if (ee == hipSuccess && g) (void)hipGraphDestroy(g);
        const hipError_t e = le != hipSuccess ? le : (ke != hipSuccess ? ke : ee);
        return fail(ENSC_ERR_HIP, std::string("step graph capture: ") + hipGetErrorString(e));
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(ENSC_ERR_HIP, std::string("step graph instantiate: ") + hipGetErrorString(ie));
    }
    E->graph = g;
    E->graph_exec = x;
    E->graph_valid = true;
    return ENSC_OK;
}

ensc_status enqueue_ticks(ensc_engine* E, int64_t n) {
    ensc_status st;
    if ((E->dirty || E->layout_scratch != E->in_scratch) && (st = relayout(E))) return st;
    if (E->timing) {
        const size_t need = 4 * (size_t)n + 1;
        while (E->evs.size() < need) {
            hipEvent_t ev;
            HIPCHK(hipEventCreate(&ev));
            E->evs.push_back(ev);
        }
        HIPCHK(hipEventRecord(E->evs[0], E->stream));
        for (int64_t i = 0; i < n; ++i)
            for (int p = 0; p < 4; ++p) {
                HIPCHK(launch_phase(E, p));
                HIPCHK(hipEventRecord(E->evs[4 * i + p + 1], E->stream));
            }
        E->pending_ticks = n;
    } else {
        int64_t i = 0;
        // graph instantiation allocates: the allocation-free entry point
        // launches plainly
        if (E->graph_ticks > 0 && n >= E->graph_ticks && !E->in_scratch) {
            if (!E->graph_valid && (st = build_graph(E))) return st;
            for (; i + E->graph_ticks <= n; i += E->graph_ticks) HIPCHK(hipGraphLaunch(E->graph_exec, E->stream));
        }
        HIPCHK(launch_ticks(E, n - i));
        E->pending_ticks = 0;
    }
    HIPCHK(hipGetLastError());
    return ENSC_OK;
}

ensc_status finish_timing(ensc_engine* E) {
    if (!E->timing || E->pending_ticks == 0) return ENSC_OK;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t i = 0; i < E->pending_ticks; ++i)
        for (int p = 0; p < 4; ++p) {
            float a = 0.f;
            HIPCHK(hipEventElapsedTime(&a, E->evs[4 * i + p], E->evs[4 * i + p + 1]));
            acc[p] += a;
        }
    for (int p = 0; p < 4; ++p) E->phase_ms[p] = acc[p] / E->pending_ticks;
    E->elem_ms = E->phase_ms[0];
    E->node_ms = E->phase_ms[2];
    E->launches = E->pending_ticks;
    E->pending_ticks = 0;
    return ENSC_OK;
}

ensc_status sync(ensc_engine* E) {
    HIPCHK(hipStreamSynchronize(E->stream));
    return finish_timing(E);
}

// Blocks of the pipelined pass: `per_cu` per CU (80 KiB of LDS each, so at
// most two), a multiple of the XCD count.
void set_pipe_grid(ensc_engine* E, int per_cu) {
    E->pipe_per_cu = per_cu;
    E->pipe_grid = (uint32_t)std::max(8, (E->ncu * per_cu + 7) / 8 * 8);
}

ensc_status engine_new(const ensc_options* opts, std::unique_ptr<ensc_engine>& E) {
    ensc_options o{};
    if (opts) o = *opts;
    if (o.precision != ENSC_FP64 && o.precision != ENSC_MIXED) return fail(ENSC_ERR_ARG, "unknown precision");
    if (ensc_device_count() <= 0) return fail(ENSC_ERR_NODEV, "no HIP device visible");
    E.reset(new ensc_engine());
    E->device = o.device;
    E->precision = o.precision;
    E->reorder = o.reorder;
    E->staging_opt = o.staging;
    auto okblk = [](int b) { return b == 64 || b == 128 || b == 256; };
    // defaults from block-size / occupancy sweeps on MI355X (the 1M-element
    // J2 cube):
    // 256 elements per block (two waves per SIMD) for the mixed and the
    // default FP64 bodies; 128 for FP64 in the reference's order (its body
    // needs > 256 registers: one wave per SIMD)
    const bool ref_order = o.precision == ENSC_FP64 && o.fp64_order == 1;
    E->blk_e = okblk(o.block_elem) ? o.block_elem : (ref_order ? 128 : 256);
    E->blk_n = okblk(o.block_node) ? o.block_node : 256;
    if (o.waves < 0 || o.waves > 3) return fail(ENSC_ERR_ARG, "waves must be 0 (default), 1, 2 or 3");
    if (o.fp64_jobs < 0 || o.fp64_jobs > 2) return fail(ENSC_ERR_ARG, "fp64_jobs must be 0 (auto), 1 or 2");
    E->fp64_rule = o.fp64_jobs;
    if (o.fp64_order < 0 || o.fp64_order > 1) return fail(ENSC_ERR_ARG, "fp64_order must be 0 or 1");
    E->fp64_order = o.fp64_order;
    E->waves = o.waves ? o.waves : (o.precision == ENSC_FP64 ? 1 : 2);
    // FP64 jobs: the engine's occupancy target in an FP64 engine, one wave
    // per SIMD in a mixed one (the FP64 body needs more than 128 registers)
    E->waves64 = o.precision == ENSC_FP64 ? E->waves : 1;
    HIPCHK(hipSetDevice(E->device));
    HIPCHK(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&E->side, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&E->ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&E->ev_join, hipEventDisableTiming));
    HIPCHK(hipDeviceGetAttribute(&E->ncu, hipDeviceAttributeMultiprocessorCount, E->device));
    set_pipe_grid(E.get(), 2);
    E->pipe_opt = o.staging == 0;
    return ENSC_OK;
}

// Upload a prepared job into a free slot.  The engine must be idle.
ensc_status add_prepared(ensc_engine* E, std::unique_ptr<Job> J, const ensc_options* opts, uint32_t* slot_out) {
    ensc_status st = upload_job(*J);
    if (st) return st;
    if (opts && opts->n_probe) {
        if (opts->n_probe > MAX_PROBE) return fail(ENSC_ERR_ARG, "at most 8 probe nodes");
        for (uint32_t i = 0; i < opts->n_probe; ++i) {
            if (opts->probe_nodes[i] >= J->model.Nn) return fail(ENSC_ERR_ARG, "probe node out of range");
            J->dev.probe[i] = J->n_old2new[opts->probe_nodes[i]];
        }
        J->dev.n_probe = (int32_t)opts->n_probe;
    }
    if (opts && opts->hist_cap > 0) {
        const size_t bytes = (size_t)opts->hist_cap * (ENSC_H_BASE + 3 * J->dev.n_probe) * sizeof(double);
        J->d_hist = static_cast<double*>(J->arena.alloc(bytes));
        if (!J->d_hist) return fail(ENSC_ERR_MEM, "history allocation failed");
        HIPCHK(hipMemset(J->d_hist, 0, bytes));
        J->dev.hist = J->d_hist;
        J->dev.hist_cap = opts->hist_cap;
        J->hist_cap = opts->hist_cap;
    }
    uint32_t slot = 0;
    while (slot < E->slots.size() && E->slots[slot]) ++slot;
    if (slot == E->slots.size()) E->slots.emplace_back();
    if ((st = grow_ctl(E, slot + 1))) return st;
    if ((st = read_ctl(E))) return st;
    Ctl& c = E->h_ctl[slot];
    std::memset(&c, 0, sizeof(Ctl));
    c.total_time = J->model.total_time;
    c.dts_min = std::numeric_limits<double>::infinity();
    c.step_limit = std::numeric_limits<long long>::max();
    c.ke0 = J->ke0;
    c.ke = J->ke0;
    c.grav_amp = 1.0;
    J->dev.ctl = E->d_ctl + slot;
    E->slots[slot] = std::move(J);
    E->dirty = true;
    if ((st = write_ctl(E))) return st;
    *slot_out = slot;
    return ENSC_OK;
}

// Build one job into a free slot.  The engine must be idle.
ensc_status add_job(ensc_engine* E, const ensc_model_desc* model, const ensc_options* opts, uint32_t* slot_out) {
    std::unique_ptr<Job> J(new Job());
    std::string err;
    if (!copy_model(model, J->model, err)) return fail(ENSC_ERR_ARG, err);
    ensc_status st = prepare_job(*J, E->precision, E->reorder, E->blk_e, E->blk_n, E->fp64_rule);
    if (st) return st;
    return add_prepared(E, std::move(J), opts, slot_out);
}

// setCurrentStep for one job: host inheritance, BC arrays, device BC
// application and the new step end time (enqueued, stream ordered).
ensc_status job_switch_step(ensc_engine* E, uint32_t slot, uint32_t si) {
    Job& J = *E->slots[slot];
    if (!job_set_step(J, si)) return ENSC_OK;
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
    // surface pressures of the step (fresh arrays: switches are rare)
    {
        std::vector<uint32_t> fnodes, nptr, nlist;
        std::vector<double> fval;
        std::vector<int32_t> famp;
        job_dsload_arrays(J, fnodes, fval, famp, nptr, nlist);
        const uint32_t nf = (uint32_t)fval.size();
        J.dev.n_dsf = nf;
        if (nf) {
            uint32_t* d_fn = static_cast<uint32_t*>(J.arena.alloc(fnodes.size() * 4));
            double* d_fv = static_cast<double*>(J.arena.alloc(fval.size() * 8));
            int32_t* d_fa = static_cast<int32_t*>(J.arena.alloc(famp.size() * 4));
            double* d_ff = static_cast<double*>(J.arena.alloc((size_t)nf * 3 * 8));
            uint32_t* d_np = static_cast<uint32_t*>(J.arena.alloc(nptr.size() * 4));
            uint32_t* d_nl = static_cast<uint32_t*>(J.arena.alloc(nlist.size() * 4));
            if (!d_fn || !d_fv || !d_fa || !d_ff || !d_np || !d_nl) return fail(ENSC_ERR_MEM, "*Dsload arrays: allocation failed");
            HIPCHK(hipMemcpy(d_fn, fnodes.data(), fnodes.size() * 4, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(d_fv, fval.data(), fval.size() * 8, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(d_fa, famp.data(), famp.size() * 4, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(d_np, nptr.data(), nptr.size() * 4, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(d_nl, nlist.data(), nlist.size() * 4, hipMemcpyHostToDevice));
            J.dev.dsf_nodes = d_fn;
            J.dev.dsf_val = d_fv;
            J.dev.dsf_amp = d_fa;
            J.dev.dsf_force = d_ff;
            J.dev.ds_ptr = d_np;
            J.dev.ds_list = d_nl;
        }
    }
    const GravState& g = J.grav[si];
    J.dev.grav_on = g.on;
    J.dev.grav_amp = g.amp;
    J.dev.grav_g = g.g;
    for (int d = 0; d < 3; ++d) J.dev.grav_dir[d] = g.dir[d];
    // the BC kernel reads this job's table entry: upload it to a scratch slot
    if (!J.d_self) {
        J.d_self = static_cast<JobDev*>(J.arena.alloc(sizeof(JobDev)));
        if (!J.d_self) return fail(ENSC_ERR_MEM, "job table allocation failed");
    }
    J.self_copy = J.dev;
    HIPCHK(hipMemcpyAsync(J.d_self, &J.self_copy, sizeof(JobDev), hipMemcpyHostToDevice, E->stream));
    const uint32_t n = (uint32_t)nodes.size();
    hipLaunchKernelGGL(k_apply_bc, dim3(1), dim3(APPLY_BC_BLK), 0, E->stream, J.d_self, 0u,
                       (const uint32_t*)J.d_bcnodes, n, J.model.steps[si].period, 1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(E->stream));  // host vectors above go out of scope
    J.pending_step = false;
    E->dirty = true;  // table entries changed (gravity, BC arrays)
    return ENSC_OK;
}

ensc_status do_init(ensc_engine* E) {
    ensc_status st;
    if ((st = sync(E)) || (st = read_ctl(E))) return st;
    bool any = false;
    for (uint32_t s = 0; s < E->slots.size(); ++s) {
        Job* J = E->slots[s].get();
        if (!J || J->inited) continue;
        E->h_ctl[s].init_pending = 1;
        J->inited = true;
        any = true;
    }
    if (!any) return ENSC_OK;  // init() runs once (exDyna3D.cpp:109-116)
    if ((st = write_ctl(E))) return st;
    E->dirty = true;
    const bool t = E->timing;
    E->timing = false;
    st = enqueue_ticks(E, 1);
    E->timing = t;
    if (st) return st;
    return sync(E);
}

ensc_status do_set_step(ensc_engine* E, uint32_t si) {
    ensc_status st;
    if ((st = sync(E))) return st;
    for (uint32_t s = 0; s < E->slots.size(); ++s) {
        Job* J = E->slots[s].get();
        if (!J) continue;
        if (!J->inited) return fail(ENSC_ERR_STATE, "ensc_engine_init must run before set_step");
        if ((st = job_switch_step(E, s, si))) return st;
    }
    if ((st = read_ctl(E))) return st;
    for (uint32_t s = 0; s < E->slots.size(); ++s)
        if (E->slots[s]) E->h_ctl[s].step_index = (int32_t)E->slots[s]->cur_step;
    if ((st = write_ctl(E))) return st;
    E->dirty = true;
    return ENSC_OK;
}

// Run ticks until every job has reached its step end (or max_steps more
// steps, or an error flag).  Finished jobs drop out of the layout.
// ---- output snapshots (single-job engines) ----
bool output_engine(const ensc_engine* E) { return E->slots.size() == 1 && E->slots[0] && E->h_ctl[0].out_on; }

// The reference's rule for the output interval when none is given
// (calculate_time_interval, exDyna3D.cpp:1947-1977).
double reference_interval(double total) {
    double ti;
    if (total <= 0.0) {
        ti = 1e-2;
    } else {
        ti = 1.0 / 50.0;
        const int frames = static_cast<int>(total / ti) + 1;
        if (frames < 10)
            ti = total / 10.0;
        else if (frames > 1000)
            ti = total / 1000.0;
    }
    if (ti > total && total > 0.0) ti = total;
    return ti;
}

// Fields of the oldest undrained snapshot of job 0 into host buffers in the
// caller's numbering.  node: [Nn][12] (x, a, v, u); post: [Ne][13] (k_post).
ensc_status snapshot_fields(ensc_engine* E, std::vector<double>* node, std::vector<double>* post, double* time) {
    Job& J = *E->slots[0];
    const Ctl& c = E->h_ctl[0];
    if (!c.out_on || c.snap_head - c.snap_tail <= 0) return fail(ENSC_ERR_STATE, "no output snapshot is pending");
    const int32_t slot = c.snap_tail % c.snap_cap;
    const size_t Nn = J.model.Nn, Ne = J.model.Ne;
    if (time) HIPCHK(hipMemcpy(time, J.dev.snap_time + slot, sizeof(double), hipMemcpyDeviceToHost));
    if (node) {
        std::vector<double> t(Nn * 12);
        HIPCHK(hipMemcpy(t.data(), J.dev.snap_node + (size_t)slot * Nn * 12, t.size() * sizeof(double),
                         hipMemcpyDeviceToHost));
        node->resize(Nn * 12);
        for (size_t i = 0; i < Nn; ++i)
            std::memcpy(node->data() + 12 * (size_t)J.n_new2old[i], t.data() + 12 * i, 12 * sizeof(double));
    }
    if (post) {
        if (!J.d_self) {
            J.d_self = static_cast<JobDev*>(J.arena.alloc(sizeof(JobDev)));
            if (!J.d_self) return fail(ENSC_ERR_MEM, "job table allocation failed");
        }
        J.self_copy = J.dev;
        HIPCHK(hipMemcpyAsync(J.d_self, &J.self_copy, sizeof(JobDev), hipMemcpyHostToDevice, E->stream));
        const uint32_t blocks = (uint32_t)((Ne + 255) / 256);
        if (J.precision == ENSC_MIXED)
            hipLaunchKernelGGL(k_post<true>, dim3(blocks), dim3(256), 0, E->stream, J.d_self, slot, J.d_post);
        else
            hipLaunchKernelGGL(k_post<false>, dim3(blocks), dim3(256), 0, E->stream, J.d_self, slot, J.d_post);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(E->stream));
        std::vector<double> t(Ne * 13);
        HIPCHK(hipMemcpy(t.data(), J.d_post, t.size() * sizeof(double), hipMemcpyDeviceToHost));
        post->resize(Ne * 13);
        for (size_t i = 0; i < Ne; ++i)
            std::memcpy(post->data() + 13 * (size_t)J.e_new2old[i], t.data() + 13 * i, 13 * sizeof(double));
    }
    return ENSC_OK;
}

// output_results' file for the oldest snapshot: the point arrays a, force,
// v, u, mass, the cell arrays vonMises, Stress, Green-Lagrange Strain
// (exDyna3D.cpp:228-309), the snapshot's positions and time.
ensc_status write_snapshot(ensc_engine* E, const char* path) {
    Job& J = *E->slots[0];
    std::vector<double> node, post;
    double time = 0.0;
    ensc_status st;
    if ((st = snapshot_fields(E, &node, &post, &time))) return st;
    const size_t Nn = J.model.Nn, Ne = J.model.Ne;
    std::vector<double> x(3 * Nn), a(3 * Nn), v(3 * Nn), u(3 * Nn), f(3 * Nn, 0.0), m(Nn);
    std::vector<double> pos(Nn * POS_REC);
    HIPCHK(hipMemcpy(pos.data(), J.dev.pos, pos.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < Nn; ++i)
        for (int d = 0; d < 3; ++d) {
            x[3 * i + d] = node[12 * i + d];
            a[3 * i + d] = node[12 * i + 3 + d];
            v[3 * i + d] = node[12 * i + 6 + d];
            u[3 * i + d] = node[12 * i + 9 + d];
        }
    for (size_t i = 0; i < Nn; ++i) {
        const double im = pos[i * POS_REC + 3];
        m[J.n_new2old[i]] = im != 0.0 ? 1.0 / im : 0.0;
    }
    std::vector<double> vm(Ne), sg(6 * Ne), gl(6 * Ne);
    for (size_t e = 0; e < Ne; ++e) {
        vm[e] = post[13 * e];
        for (int q = 0; q < 6; ++q) {
            sg[6 * e + q] = post[13 * e + 1 + q];
            gl[6 * e + q] = post[13 * e + 7 + q];
        }
    }
    // solution_f is never written by the reference (zero, exDyna3D.cpp:179)
    const ensc_vtu_array pd[5] = {{"a", 3, 0, a.data()},
                                  {"force", 3, 0, f.data()},
                                  {"v", 3, 0, v.data()},
                                  {"u", 3, 0, u.data()},
                                  {"mass", 1, 0, m.data()}};
    const ensc_vtu_array cd[3] = {
        {"vonMises", 1, 0, vm.data()}, {"Stress", 6, 0, sg.data()}, {"Green-Lagrange Strain", 6, 0, gl.data()}};
    return ensc_vtu_write(path, (uint32_t)Nn, x.data(), (uint32_t)Ne, J.model.conn.data(), time, pd, 5, cd, 3);
}

ensc_status pop_snapshot(ensc_engine* E) {
    Ctl& c = E->h_ctl[0];
    if (!c.out_on || c.snap_head - c.snap_tail <= 0) return fail(ENSC_ERR_STATE, "no output snapshot is pending");
    c.snap_tail += 1;
    return write_ctl(E);
}

// With an output directory: write every pending snapshot as
// <dir>/solution_<n>.vtu and drop it.  Without one: *full tells whether the
// ring is full (the job then waits for the caller).
ensc_status drain_outputs(ensc_engine* E, bool* full) {
    ensc_status st;
    Ctl& c = E->h_ctl[0];
    if (!E->out_dir.empty()) {
        while (c.snap_head - c.snap_tail > 0) {
            const std::string path = E->out_dir + "/solution_" + std::to_string(E->out_count) + ".vtu";
            if ((st = write_snapshot(E, path.c_str()))) return st;
            ++E->out_count;
            if ((st = pop_snapshot(E))) return st;
        }
    }
    *full = c.snap_head - c.snap_tail >= c.snap_cap;
    return ENSC_OK;
}

ensc_status do_advance(ensc_engine* E, int64_t max_steps, int32_t* done) {
    ensc_status st;
    if ((st = sync(E)) || (st = read_ctl(E))) return st;
    for (uint32_t s = 0; s < E->slots.size(); ++s)
        if (E->slots[s])
            E->h_ctl[s].step_limit =
                max_steps > 0 ? E->h_ctl[s].nsteps + max_steps : std::numeric_limits<long long>::max();
    if ((st = write_ctl(E))) return st;
    E->dirty = true;
    int64_t chunk = std::max<int64_t>(4, E->graph_ticks);
    size_t n_live_blocks = 0;
    for (;;) {
        bool all = true;
        for (uint32_t s = 0; s < E->slots.size(); ++s) {
            Job* J = E->slots[s].get();
            if (J) all = all && job_done(E->h_ctl[s]);
        }
        if (all) break;
        // drop finished jobs from the launch once they hold a quarter of its
        // element blocks
        size_t done_blocks = 0;
        for (uint32_t s : E->live)
            if (job_done(E->h_ctl[s])) done_blocks += E->slots[s]->dev.nblk_e;
        if (n_live_blocks && done_blocks * 4 > n_live_blocks) E->dirty = true;
        if (E->dirty) {
            if ((st = relayout(E))) return st;
            n_live_blocks = (size_t)E->nblk_pipe + E->nblk_e + E->nblk_g + E->nblk_e64 + E->nblk_w64 + E->nblk_g64;
        }
        int64_t n = chunk;
        if (output_engine(E)) {
            // output snapshots: drain the ring, and size the chunk so the free
            // slots last (about t_int / dt_i steps per output)
            bool full = false;
            if ((st = drain_outputs(E, &full))) return st;
            if (full) break;  // the caller drains (no output directory)
            const Ctl& c = E->h_ctl[0];
            const double per = c.dt_i > 0.0 ? c.t_int / c.dt_i : 1.0;
            const double fit = (double)(c.snap_cap - (c.snap_head - c.snap_tail)) * std::max(per, 1.0);
            n = std::max<int64_t>(1, std::min<int64_t>(n, (int64_t)fit));
        }
        const bool t = E->timing;
        E->timing = false;
        st = enqueue_ticks(E, n);
        E->timing = t;
        if (st) return st;
        if ((st = read_ctl(E))) return st;
        chunk = std::min<int64_t>(chunk * 2, std::max<int64_t>(256, E->graph_ticks));
        if (E->graph_ticks > 0) chunk = (chunk / E->graph_ticks) * E->graph_ticks;
    }
    if (output_engine(E)) {
        bool full = false;
        if ((st = drain_outputs(E, &full))) return st;
    }
    bool step_end = true, errs = false;
    for (uint32_t s = 0; s < E->slots.size(); ++s) {
        if (!E->slots[s]) continue;
        step_end = step_end && !(E->h_ctl[s].time < E->h_ctl[s].step_end);
        errs = errs || E->h_ctl[s].err != 0;
        E->h_ctl[s].step_limit = std::numeric_limits<long long>::max();
    }
    if ((st = write_ctl(E))) return st;
    if (done) *done = step_end ? 1 : 0;
    if (errs) return fail(ENSC_ERR_NUMERIC, "a job raised a device error flag (see ENSC_S_ERR)");
    return ENSC_OK;
}

void fill_scalars(const Ctl& c, double* out) {
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
}

ensc_status download(ensc_engine* E, Job& J, const std::string& f, double* out, int64_t count) {
    const size_t Nn = J.model.Nn, Ne = J.model.Ne, es = J.ne_s;
    const bool mixed = J.precision == ENSC_MIXED;
    if (f == "u" || f == "v" || f == "x") {
        if (count < (int64_t)(3 * Nn)) return fail(ENSC_ERR_ARG, "buffer too small");
        std::vector<double> pos(Nn * POS_REC), vu(Nn * VU_REC);
        HIPCHK(hipMemcpy(pos.data(), J.dev.pos, pos.size() * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(vu.data(), J.dev.vu, vu.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < Nn; ++i) {
            const size_t o = J.n_new2old[i];
            for (int d = 0; d < 3; ++d) {
                double x;
                if (f == "v")
                    x = vu[i * VU_REC + d];
                else if (f == "u")
                    x = vu[i * VU_REC + 3 + d];
                else  // the mixed engine keeps the reference position: x = X + u
                    x = mixed ? pos[i * POS_REC + d] + vu[i * VU_REC + 3 + d] : pos[i * POS_REC + d];
                out[3 * o + d] = x;
            }
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

}  // namespace

// Internal entry points used by the serving loop (server.cpp).
namespace ensc {
ensc_status engine_new_public(const ensc_options* o, ensc_engine** out) {
    std::unique_ptr<ensc_engine> E;
    ensc_status st = engine_new(o, E);
    if (st) return st;
    *out = E.release();
    return ENSC_OK;
}
ensc_status engine_add_job(ensc_engine* E, const ensc_model_desc* m, uint32_t* slot) {
    ensc_status st = sync(E);
    if (st) return st;
    return add_job(E, m, nullptr, slot);
}
// Host preparation of a job for this engine's settings; touches no engine
// state, so a worker thread may run it while the engine steps.
ensc_status engine_prepare_job(const ensc_engine* E, ModelCopy&& model, std::unique_ptr<Job>& out) {
    std::unique_ptr<Job> J(new Job());
    J->model = std::move(model);
    ensc_status st = prepare_job(*J, E->precision, E->reorder, E->blk_e, E->blk_n, E->fp64_rule);
    if (st) return st;
    out = std::move(J);
    return ENSC_OK;
}
ensc_status engine_add_prepared(ensc_engine* E, std::unique_ptr<Job> J, uint32_t* slot) {
    ensc_status st = sync(E);
    if (st) return st;
    return add_prepared(E, std::move(J), nullptr, slot);
}
// Stream-ordered start of a newly added job: its init pass runs in the next
// tick together with every other live job; its step 1 is switched in after.
ensc_status engine_start_jobs(ensc_engine* E, const std::vector<uint32_t>& slots) {
    ensc_status st;
    if ((st = sync(E)) || (st = read_ctl(E))) return st;
    for (uint32_t s : slots) {
        Job* J = job_at(E, s);
        if (!J) return fail(ENSC_ERR_ARG, "bad job id");
        E->h_ctl[s].init_pending = 1;
        J->inited = true;
        J->pending_step = true;
    }
    if ((st = write_ctl(E))) return st;
    E->dirty = true;
    // init pass for the new jobs (others step normally in the same tick)
    if ((st = enqueue_ticks(E, 1))) return st;
    if ((st = sync(E))) return st;
    for (uint32_t s : slots)
        if ((st = job_switch_step(E, s, 1))) return st;
    if ((st = read_ctl(E))) return st;
    for (uint32_t s : slots) E->h_ctl[s].step_index = 1;
    if ((st = write_ctl(E))) return st;
    E->dirty = true;
    return ENSC_OK;
}
ensc_status engine_ticks(ensc_engine* E, int64_t n) { return enqueue_ticks(E, n); }
// Jobs that reached the end of their current *Step: those with a later step
// switch to it (setCurrentStep, as exDyna3D::run's step loop does,
// exDyna3D.cpp:77-99) and keep running; the others (and jobs with an error
// flag) are reported finished.
ensc_status engine_poll(ensc_engine* E, std::vector<uint32_t>& finished) {
    ensc_status st;
    if ((st = sync(E)) || (st = read_ctl(E))) return st;
    finished.clear();
    std::vector<uint32_t> next;
    for (uint32_t s = 0; s < E->slots.size(); ++s) {
        Job* J = E->slots[s].get();
        if (!J || !J->inited || J->pending_step) continue;
        const Ctl& c = E->h_ctl[s];
        if (c.init_pending) continue;
        if (!(c.time < c.step_end) || c.err) {
            if (!c.err && J->cur_step + 1 < J->model.steps.size())
                next.push_back(s);
            else
                finished.push_back(s);
        }
    }
    if (next.empty()) return ENSC_OK;
    for (uint32_t s : next)
        if ((st = job_switch_step(E, s, E->slots[s]->cur_step + 1))) return st;
    if ((st = read_ctl(E))) return st;
    for (uint32_t s : next) E->h_ctl[s].step_index = (int32_t)E->slots[s]->cur_step;
    if ((st = write_ctl(E))) return st;
    E->dirty = true;
    return ENSC_OK;
}
ensc_status engine_job_scalars(ensc_engine* E, uint32_t s, double* out) {
    ensc_status st;
    if ((st = sync(E)) || (st = read_ctl(E))) return st;
    if (!job_at(E, s)) return fail(ENSC_ERR_ARG, "bad job id");
    fill_scalars(E->h_ctl[s], out);
    return ENSC_OK;
}
ensc_status engine_job_download(ensc_engine* E, uint32_t s, const char* f, double* out, int64_t count) {
    ensc_status st = sync(E);
    if (st) return st;
    Job* J = job_at(E, s);
    if (!J) return fail(ENSC_ERR_ARG, "bad job id");
    return download(E, *J, f, out, count);
}
ensc_status engine_remove_job(ensc_engine* E, uint32_t s) {
    ensc_status st = sync(E);
    if (st) return st;
    if (!job_at(E, s)) return fail(ENSC_ERR_ARG, "bad job id");
    E->slots[s].reset();
    E->dirty = true;
    return ENSC_OK;
}
int64_t engine_job_bytes(ensc_engine* E, uint32_t s) {
    Job* J = job_at(E, s);
    return J ? (int64_t)J->arena.total : 0;
}
uint32_t engine_job_blocks(ensc_engine* E, uint32_t s) {
    Job* J = job_at(E, s);
    return J ? J->dev.nblk_e : 0;
}
// Make the engine's device current on the calling thread (allocations and
// plain copies use the current device).
ensc_status engine_bind_device(ensc_engine* E) {
    HIPCHK(hipSetDevice(E->device));
    return ENSC_OK;
}
ensc_status engine_set_graph(ensc_engine* E, int32_t ticks) {
    E->graph_ticks = std::max(0, ticks);
    E->graph_valid = false;
    return ENSC_OK;
}

// Autotuner.  Trial runs advance the jobs, so the state a tick changes is
// copied aside first and put back after every trial: per live job v and u,
// the u compensation (mixed) or the positions (FP64), sigma, QIA, plastic
// strain and reactions, and every job's scalars.  Between runs of ticks the
// nodal update is always closed (each run ends with the node finalize), so
// no block partials carry over, and the history and output rings are
// rewritten as the restored steps run again.
struct SavedState {
    std::vector<std::pair<void*, size_t>> regions;
    char* buf = nullptr;
    ~SavedState() {
        if (buf) (void)hipFree(buf);
    }
};
ensc_status save_state(ensc_engine* E, SavedState& S) {
    for (uint32_t s : E->live) {
        const Job& J = *E->slots[s];
        const size_t R = J.precision == ENSC_MIXED ? sizeof(float) : sizeof(double);
        const size_t nn = J.model.Nn, es = J.ne_s;
        S.regions.push_back({J.dev.vu, nn * VU_REC * sizeof(double)});
        if (J.dev.ucomp) S.regions.push_back({J.dev.ucomp, nn * sizeof(uint32_t)});
        if (J.precision != ENSC_MIXED) S.regions.push_back({J.dev.pos, nn * POS_REC * sizeof(double)});
        S.regions.push_back({J.dev.sigma, 6 * es * R});
        S.regions.push_back({J.dev.qia, 12 * es * R});
        S.regions.push_back({J.dev.epsp, es * R});
        S.regions.push_back({J.dev.rf, nn * 3 * sizeof(double)});
    }
    S.regions.push_back({E->d_ctl, (size_t)E->ctl_cap * sizeof(Ctl)});
    size_t total = 0;
    for (const auto& r : S.regions) total += round_up(r.second, 256);
    if (hipMalloc(&S.buf, total) != hipSuccess) {
        S.buf = nullptr;
        return fail(ENSC_ERR_MEM, "autotune: no device memory for the saved state (" + std::to_string(total) + " bytes)");
    }
    size_t off = 0;
    for (const auto& r : S.regions) {
        HIPCHK(hipMemcpyAsync(S.buf + off, r.first, r.second, hipMemcpyDeviceToDevice, E->stream));
        off += round_up(r.second, 256);
    }
    HIPCHK(hipStreamSynchronize(E->stream));
    return ENSC_OK;
}
ensc_status restore_state(ensc_engine* E, const SavedState& S) {
    size_t off = 0;
    for (const auto& r : S.regions) {
        HIPCHK(hipMemcpyAsync(r.first, S.buf + off, r.second, hipMemcpyDeviceToDevice, E->stream));
        off += round_up(r.second, 256);
    }
    HIPCHK(hipStreamSynchronize(E->stream));
    return read_ctl(E);
}
void set_node_block(ensc_engine* E, int blk) {
    E->blk_n = blk;
    for (auto& j : E->slots)
        if (j) j->dev.nblk_n = (uint32_t)((j->model.Nn + blk - 1) / blk);
    E->dirty = true;
}
// Mean milliseconds per tick of `ticks` ticks (after two untimed ones; the
// layout and the step graph of the candidate are built before the clock
// starts).
ensc_status time_ticks(ensc_engine* E, int32_t ticks, double* ms) {
    ensc_status st;
    if ((st = enqueue_ticks(E, 2)) || (st = sync(E))) return st;
    if (E->dirty && (st = relayout(E))) return st;
    if (E->graph_ticks > 0 && ticks >= E->graph_ticks && !E->graph_valid && (st = build_graph(E))) return st;
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    float t = 0.f;
    hipError_t e = hipEventRecord(a, E->stream);
    if (e == hipSuccess && !(st = enqueue_ticks(E, ticks))) {
        e = hipEventRecord(b, E->stream);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (st) return st;
    if (e != hipSuccess) return fail(ENSC_ERR_HIP, std::string("autotune timing: ") + hipGetErrorString(e));
    *ms = (double)t / ticks;
    return ENSC_OK;
}
// Coordinate search over the launch choices that leave every job's tables
// and element arithmetic unchanged: blocks per CU of the pipelined pass,
// waves per SIMD of the staged element kernel, and the nodal update's block
// size (which regroups only the kinetic-energy and external-work block
// sums).  Keeps the fastest; the jobs' state is as before the call.
ensc_status autotune(ensc_engine* E, int32_t ticks, int64_t* chosen) {
    ensc_status st;
    if ((st = sync(E))) return st;
    if (E->dirty && (st = relayout(E))) return st;
    if (E->live.empty()) return fail(ENSC_ERR_STATE, "autotune: no job to step");
    SavedState S;
    if ((st = save_state(E, S))) return st;
    const bool any_pipe = E->nblk_pipe > 0, any_staged = E->nblk_e + E->nblk_g + E->nblk_e64 + E->nblk_w64 + E->nblk_g64 > 0;
    auto trial = [&](double* ms) -> ensc_status {
        ensc_status s2 = time_ticks(E, ticks, ms);
        ensc_status s3 = restore_state(E, S);
        return s2 ? s2 : s3;
    };
    double best = 0.0;
    if ((st = trial(&best))) return st;
    if (any_pipe) {
        const int cur = E->pipe_per_cu;
        int pick = cur;
        for (int p : {1, 2}) {
            if (p == cur) continue;
            set_pipe_grid(E, p);
            E->graph_valid = false;
            double ms = 0.0;
            if ((st = trial(&ms))) return st;
            if (ms < best) best = ms, pick = p;
        }
        set_pipe_grid(E, pick);
        E->graph_valid = false;
    }
    if (any_staged) {
        const int cur = E->waves;
        int pick = cur;
        for (int w : {1, 2, 3}) {
            if (w == cur) continue;
            E->waves = w;
            E->dirty = true;
            double ms = 0.0;
            if ((st = trial(&ms))) return st;
            if (ms < best) best = ms, pick = w;
        }
        E->waves = pick;
        E->dirty = true;
    }
    {
        const int cur = E->blk_n;
        int pick = cur;
        for (int b : {64, 128, 256}) {
            if (b == cur) continue;
            set_node_block(E, b);
            double ms = 0.0;
            if ((st = trial(&ms))) return st;
            if (ms < best) best = ms, pick = b;
        }
        set_node_block(E, pick);
    }
    if (E->dirty && (st = relayout(E))) return st;
    if (chosen) {
        chosen[0] = E->waves;
        chosen[1] = E->blk_n;
        chosen[2] = E->pipe_per_cu;
        chosen[3] = (int64_t)(best * 1e6);  // ns per tick of the choice
    }
    return ENSC_OK;
}
}  // namespace ensc

// ---- allocation-free stepping (ensc_engine_advance_static) ----
// The job table, the block maps and the block partials of the call live in
// this file-scope device array instead of engine buffers.  Each GPU holds its
// own copy for the life of the process; in host code its name is only a
// stand-in, so kernels and copies get its device address from
// hipGetSymbolAddress, taken once per device on the first call there.
static __device__ __align__(256) unsigned char scratch[STATIC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
char* g_scratch[MAX_DEVICES] = {};

// The scratch array's address on `device`, which must be current.
ensc_status scratch_on(int device, char** out) {
    if (device < 0 || device >= MAX_DEVICES) return fail(ENSC_ERR_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(ENSC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[device] = static_cast<char*>(p);
    }
    *out = g_scratch[device];
    return ENSC_OK;
}
}  // namespace

extern "C" {

const char* ensc_last_error(void) { return g_last_error.c_str(); }

int32_t ensc_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

ensc_status ensc_copy_bandwidth(int32_t device, int64_t bytes, int32_t reps, double* gbs) {
    if (bytes < 16 || reps < 1 || !gbs) return fail(ENSC_ERR_ARG, "bytes >= 16, reps >= 1 and gbs required");
    if (ensc_device_count() <= 0) return fail(ENSC_ERR_NODEV, "no HIP device visible");
    HIPCHK(hipSetDevice(device));
    const size_t n = (size_t)bytes / 16;
    void *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, n * 16));
    if (hipMalloc(&b, n * 16) != hipSuccess) {
        (void)hipFree(a);
        return fail(ENSC_ERR_MEM, "copy buffer allocation failed");
    }
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ensc_status st = ENSC_OK;
    float ms = 0.f;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess || hipMemsetAsync(a, 0, n * 16, s) != hipSuccess) {
        st = fail(ENSC_ERR_HIP, "copy probe set-up failed");
    } else {
        const unsigned grid = 256 * 8 * 4;
        hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, s, (const double2*)a, (double2*)b, n);
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, s, (const double2*)a, (double2*)b, n);
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            st = fail(ENSC_ERR_HIP, "copy probe failed");
        else
            *gbs = 2.0 * (double)(n * 16) * reps / (ms * 1e-3) / 1e9;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(a);
    (void)hipFree(b);
    return st;
}

ensc_status ensc_engine_create(const ensc_model_desc* model, const ensc_options* opts, ensc_engine** out) {
    if (!out) return fail(ENSC_ERR_ARG, "out is null");
    *out = nullptr;
    return guarded([&]() -> ensc_status {
        std::unique_ptr<ensc_engine> E;
        ensc_status st = engine_new(opts, E);
        if (st) return st;
        uint32_t slot = 0;
        if ((st = add_job(E.get(), model, opts, &slot))) return st;
        E->slots[0]->keep_in_layout = true;
        if ((st = relayout(E.get()))) return st;
        *out = E.release();
        return ENSC_OK;
    });
}

ensc_status ensc_batch_create(const ensc_model_desc* const* models, uint32_t n_models, const ensc_options* opts,
                              ensc_engine** out) {
    if (!out || !models || n_models == 0) return fail(ENSC_ERR_ARG, "no models");
    *out = nullptr;
    return guarded([&]() -> ensc_status {
        std::unique_ptr<ensc_engine> E;
        ensc_status st = engine_new(opts, E);
        if (st) return st;
        if ((st = grow_ctl(E.get(), n_models))) return st;
        for (uint32_t i = 0; i < n_models; ++i) {
            uint32_t slot = 0;
            if ((st = add_job(E.get(), models[i], nullptr, &slot))) return st;
        }
        if ((st = relayout(E.get()))) return st;
        *out = E.release();
        return ENSC_OK;
    });
}

ensc_status ensc_engine_destroy(ensc_engine* E) {
    if (!E) return ENSC_OK;
    (void)hipSetDevice(E->device);
    (void)hipStreamSynchronize(E->stream);
    if (E->side) (void)hipStreamSynchronize(E->side);
    for (hipEvent_t ev : E->evs) (void)hipEventDestroy(ev);
    drop_graph(E);
    E->slots.clear();
    if (E->d_ctl) (void)hipFree(E->d_ctl);
    if (E->h_ctl) (void)hipHostFree(E->h_ctl);
    if (E->ev_fork) (void)hipEventDestroy(E->ev_fork);
    if (E->ev_join) (void)hipEventDestroy(E->ev_join);
    if (E->side) (void)hipStreamDestroy(E->side);
    if (E->stream) (void)hipStreamDestroy(E->stream);
    delete E;
    return ENSC_OK;
}

ensc_status ensc_engine_init(ensc_engine* E) {
    if (!E) return fail(ENSC_ERR_ARG, "null engine");
    HIPCHK(hipSetDevice(E->device));
    return do_init(E);
}

ensc_status ensc_engine_set_step(ensc_engine* E, uint32_t si) {
    if (!E) return fail(ENSC_ERR_ARG, "null engine");
    HIPCHK(hipSetDevice(E->device));
    return do_set_step(E, si);
}

ensc_status ensc_engine_advance(ensc_engine* E, int64_t max_steps, int32_t* done) {
    if (!E) return fail(ENSC_ERR_ARG, "null engine");
    HIPCHK(hipSetDevice(E->device));
    return do_advance(E, max_steps, done);
}

ensc_status ensc_engine_advance_static(ensc_engine* E, int64_t max_steps, int32_t* done) {
    if (!E) return fail(ENSC_ERR_ARG, "null engine");
    HIPCHK(hipSetDevice(E->device));
    return guarded([&]() -> ensc_status {
        ensc_status st;
        if ((st = sync(E))) return st;
        if (output_engine(E))
            return fail(ENSC_ERR_STATE, "ensc_engine_advance_static: output snapshots are on (they copy and write "
                                        "files); use ensc_engine_advance");
        char* base = nullptr;
        if ((st = scratch_on(E->device, &base))) return st;
        // another engine's call may have left its tables in the array: lay
        // this engine's out again, into it
        E->scratch_base = base;
        E->in_scratch = true;
        E->dirty = true;
        st = do_advance(E, max_steps, done);
        E->in_scratch = false;
        E->dirty = true;  // the next call lays out in the engine's own buffers
        return st;
    });
}

ensc_status ensc_engine_run(ensc_engine* E) {
    if (!E) return fail(ENSC_ERR_ARG, "null engine");
    HIPCHK(hipSetDevice(E->device));
    ensc_status st;
    if ((st = do_init(E))) return st;
    size_t nsteps = 0;
    for (auto& j : E->slots)
        if (j) nsteps = std::max(nsteps, j->model.steps.size());
    for (uint32_t s = 1; s < nsteps; ++s) {
        if ((st = do_set_step(E, s))) return st;
        int32_t done = 0;
        if ((st = do_advance(E, 0, &done))) return st;
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

ensc_status ensc_engine_set_graph(ensc_engine* E, int32_t ticks) {
    if (!E) return fail(ENSC_ERR_ARG, "null engine");
    HIPCHK(hipSetDevice(E->device));
    ensc_status st = sync(E);
    if (st) return st;
    return engine_set_graph(E, ticks);
}

ensc_status ensc_engine_download(ensc_engine* E, const char* field, double* out, int64_t count) {
    if (!E || !field || !out) return fail(ENSC_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(E->device));
    return engine_job_download(E, 0, field, out, count);
}

ensc_status ensc_engine_scalars(ensc_engine* E, double* out) {
    if (!E || !out) return fail(ENSC_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(E->device));
    return engine_job_scalars(E, 0, out);
}

ensc_status ensc_engine_history(ensc_engine* E, double* out, int64_t rows, int64_t* got) {
    if (!E || !got) return fail(ENSC_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(E->device));
    ensc_status st;
    if ((st = sync(E)) || (st = read_ctl(E))) return st;
    Job* J = job_at(E, 0);
    if (!J) return fail(ENSC_ERR_ARG, "no job");
    int64_t n = std::min<int64_t>(E->h_ctl[0].nsteps, J->hist_cap);
    n = std::min<int64_t>(n, rows);
    *got = n;
    if (n > 0 && out)
        HIPCHK(hipMemcpy(out, J->d_hist, (size_t)n * (ENSC_H_BASE + 3 * J->dev.n_probe) * sizeof(double),
                         hipMemcpyDeviceToHost));
    return ENSC_OK;
}

ensc_status ensc_engine_kernel_times(ensc_engine* E, double* elem_ms, double* node_ms, int64_t* launches) {
    if (!E) return fail(ENSC_ERR_ARG, "null engine");
    HIPCHK(hipSetDevice(E->device));
    ensc_status st = sync(E);
    if (st) return st;
    if (elem_ms) *elem_ms = E->elem_ms;
    if (node_ms) *node_ms = E->node_ms;
    if (launches) *launches = E->launches;
    return ENSC_OK;
}

ensc_status ensc_engine_phase_times(ensc_engine* E, double* ms4) {
    if (!E || !ms4) return fail(ENSC_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(E->device));
    ensc_status st = sync(E);
    if (st) return st;
    for (int p = 0; p < 4; ++p) ms4[p] = E->phase_ms[p];
    return ENSC_OK;
}

ensc_status ensc_engine_layout(ensc_engine* E, int64_t* out) {
    if (!E || !out) return fail(ENSC_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(E->device));
    ensc_status st = sync(E);
    if (st) return st;
    out[ENSC_L_STAGED] = E->staged ? 1 : 0;
    out[ENSC_L_PIPE_BLOCKS] = E->nblk_pipe;
    out[ENSC_L_ELEM_BLOCKS] = E->nblk_e + E->nblk_g + E->nblk_e64 + E->nblk_w64 + E->nblk_g64;
    out[ENSC_L_NODE_BLOCKS] = E->nblk_n_asm + E->nblk_n + E->nblk_n64_asm + E->nblk_n64;
    out[ENSC_L_LDS_BYTES] = (int64_t)std::max(E->lds_bytes, std::max(E->lds_bytes64, E->lds_bytes_w64));
    out[ENSC_L_ASM_JOBS] = E->n_asm;
    out[ENSC_L_FP64_JOBS] = E->n_fp64;
    return ENSC_OK;
}

ensc_status ensc_engine_set_output(ensc_engine* E, double interval, int32_t slots, const char* dir) {
    if (!E) return fail(ENSC_ERR_ARG, "null engine");
    HIPCHK(hipSetDevice(E->device));
    if (E->slots.size() != 1 || !E->slots[0]) return fail(ENSC_ERR_STATE, "output snapshots need a single-job engine");
    if (slots < 0 || slots > 64) return fail(ENSC_ERR_ARG, "slots must be 0..64");
    ensc_status st;
    if ((st = sync(E)) || (st = read_ctl(E))) return st;
    Job& J = *E->slots[0];
    Ctl& c = E->h_ctl[0];
    if (slots == 0) {
        c.out_on = 0;
        c.snap_now = 0;
        return write_ctl(E);
    }
    if ((uint32_t)slots != J.snap_slots) {
        const size_t Nn = J.model.Nn, Ne = J.model.Ne;
        J.dev.snap_node = static_cast<double*>(J.arena.alloc((size_t)slots * Nn * 12 * sizeof(double)));
        J.dev.snap_elem = static_cast<double*>(J.arena.alloc((size_t)slots * Ne * 6 * sizeof(double)));
        J.dev.snap_time = static_cast<double*>(J.arena.alloc((size_t)slots * sizeof(double)));
        if (!J.d_post) J.d_post = static_cast<double*>(J.arena.alloc(Ne * 13 * sizeof(double)));
        if (!J.dev.snap_node || !J.dev.snap_elem || !J.dev.snap_time || !J.d_post)
            return fail(ENSC_ERR_MEM, "output snapshot allocation failed");
        J.snap_slots = (uint32_t)slots;
    }
    c.out_on = 1;
    c.t_int = interval > 0.0 ? interval : reference_interval(c.total_time);
    c.t_out = c.time;  // the next step is written (the reference starts at time 0)
    c.snap_head = c.snap_tail = 0;
    c.snap_cap = slots;
    c.snap_now = 0;
    E->out_dir = dir ? dir : "";
    E->out_count = 0;
    E->dirty = true;  // the job table carries the ring's pointers
    return write_ctl(E);
}

ensc_status ensc_engine_output_pending(ensc_engine* E, int32_t* n, double* time) {
    if (!E || !n) return fail(ENSC_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(E->device));
    ensc_status st;
    if ((st = sync(E)) || (st = read_ctl(E))) return st;
    *n = output_engine(E) ? E->h_ctl[0].snap_head - E->h_ctl[0].snap_tail : 0;
    if (time) {
        *time = 0.0;
        if (*n > 0) return snapshot_fields(E, nullptr, nullptr, time);
    }
    return ENSC_OK;
}

ensc_status ensc_engine_output_field(ensc_engine* E, const char* field, double* out, int64_t count) {
    if (!E || !field || !out) return fail(ENSC_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(E->device));
    ensc_status st;
    if ((st = sync(E)) || (st = read_ctl(E))) return st;
    if (!output_engine(E)) return fail(ENSC_ERR_STATE, "output snapshots are off");
    const Job& J = *E->slots[0];
    const size_t Nn = J.model.Nn, Ne = J.model.Ne;
    const std::string f(field);
    static const char* node_f[4] = {"x", "a", "v", "u"};
    for (int q = 0; q < 4; ++q)
        if (f == node_f[q]) {
            if (count < (int64_t)(3 * Nn)) return fail(ENSC_ERR_ARG, "buffer too small");
            std::vector<double> node;
            if ((st = snapshot_fields(E, &node, nullptr, nullptr))) return st;
            for (size_t i = 0; i < Nn; ++i)
                for (int d = 0; d < 3; ++d) out[3 * i + d] = node[12 * i + 3 * q + d];
            return ENSC_OK;
        }
    int off = -1, nc = 0;
    if (f == "vonMises") {
        off = 0;
        nc = 1;
    } else if (f == "Stress") {
        off = 1;
        nc = 6;
    } else if (f == "Green-Lagrange Strain") {
        off = 7;
        nc = 6;
    } else {
        return fail(ENSC_ERR_ARG, "unknown output field '" + f + "'");
    }
    if (count < (int64_t)(nc * Ne)) return fail(ENSC_ERR_ARG, "buffer too small");
    std::vector<double> post;
    if ((st = snapshot_fields(E, nullptr, &post, nullptr))) return st;
    for (size_t e = 0; e < Ne; ++e)
        for (int q = 0; q < nc; ++q) out[(size_t)nc * e + q] = post[13 * e + off + q];
    return ENSC_OK;
}

ensc_status ensc_engine_output_write(ensc_engine* E, const char* path) {
    if (!E || !path) return fail(ENSC_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(E->device));
    ensc_status st;
    if ((st = sync(E)) || (st = read_ctl(E))) return st;
    if (!output_engine(E)) return fail(ENSC_ERR_STATE, "output snapshots are off");
    return write_snapshot(E, path);
}

ensc_status ensc_engine_output_pop(ensc_engine* E) {
    if (!E) return fail(ENSC_ERR_ARG, "null engine");
    HIPCHK(hipSetDevice(E->device));
    ensc_status st;
    if ((st = sync(E)) || (st = read_ctl(E))) return st;
    if (!output_engine(E)) return fail(ENSC_ERR_STATE, "output snapshots are off");
    return pop_snapshot(E);
}

ensc_status ensc_engine_autotune(ensc_engine* E, int32_t ticks, int64_t* chosen) {
    if (!E) return fail(ENSC_ERR_ARG, "null engine");
    if (ticks < 1 || ticks > 10000) return fail(ENSC_ERR_ARG, "autotune: ticks must be 1..10000");
    HIPCHK(hipSetDevice(E->device));
    return guarded([&] { return autotune(E, ticks, chosen); });
}

ensc_status ensc_engine_set_timing(ensc_engine* E, int32_t on) {
    if (!E) return fail(ENSC_ERR_ARG, "null engine");
    HIPCHK(hipSetDevice(E->device));
    ensc_status st = sync(E);
    if (st) return st;
    E->timing = on != 0;
    return ENSC_OK;
}

// ---- fluid-structure coupling (exDyna3D.cpp:1281-1758) ----

ensc_status ensc_engine_fsi_setup(ensc_engine* E, int32_t n_layers, double dist, int64_t* n_particles) {
    if (!E) return fail(ENSC_ERR_ARG, "null engine");
    HIPCHK(hipSetDevice(E->device));
    return guarded([&]() -> ensc_status {
        if (E->slots.size() != 1 || !job_at(E, 0))
            return fail(ENSC_ERR_ARG, "ensc_engine_fsi_setup: FSI coupling needs a single-job engine");
        if (n_layers < 0 || !(dist > 0.0) || !std::isfinite(dist))
            return fail(ENSC_ERR_ARG, "ensc_engine_fsi_setup: need n_layers >= 0 and a finite dist > 0");
        Job& J = *job_at(E, 0);
        if (!J.inited) return fail(ENSC_ERR_STATE, "ensc_engine_fsi_setup: call ensc_engine_init first");
        ensc_status st;
        if ((st = sync(E))) return st;
        const ModelCopy& M = J.model;
        // the particles are placed on the current positions
        std::vector<double> x(3 * (size_t)M.Nn);
        if ((st = download(E, J, "x", x.data(), (int64_t)x.size()))) return st;
        FsiParticles P;
        fsi_particles(M, x, n_layers, dist, P);
        const uint32_t np = (uint32_t)(P.unit.size() / 3), ns = (uint32_t)P.elems.size();
        // per particle its FSI element slot, per slot its nodes (engine numbering)
        std::vector<uint32_t> pelem(np), econn(8 * (size_t)ns);
        for (uint32_t s = 0; s < ns; ++s) {
            for (uint32_t p = P.eptr[s]; p < P.eptr[s + 1]; ++p) pelem[p] = s;
            for (int k = 0; k < 8; ++k) econn[8 * (size_t)s + k] = J.n_old2new[M.conn[8 * (size_t)P.elems[s] + k]];
        }
        // the FSI nodes and their (slot, corner) entries in slot order, which
        // is ascending original element order
        std::vector<int32_t> node_of(M.Nn, -1);
        std::vector<uint32_t> cnt(M.Nn, 0), nptr(1, 0);
        for (uint32_t n : econn) cnt[n]++;
        uint32_t nn = 0;
        for (uint32_t n = 0; n < M.Nn; ++n)
            if (cnt[n]) {
                node_of[n] = (int32_t)nn++;
                nptr.push_back(nptr.back() + cnt[n]);
            }
        std::vector<uint32_t> nent(nptr.back()), fill(nptr.begin(), nptr.end() - 1);
        for (uint32_t s = 0; s < ns; ++s)
            for (uint32_t k = 0; k < 8; ++k) nent[fill[node_of[econn[8 * (size_t)s + k]]]++] = s << 3 | k;
        JobDev& D = J.dev;
        D.n_fsi_part = 0;
        D.n_fsi_node = 0;
        E->fsi_np = E->fsi_nn = 0;
        if (np) {
            auto put = [&](const void* src, size_t bytes) -> void* {
                void* p = J.arena.alloc(std::max<size_t>(bytes, 8));
                if (p && bytes && hipMemcpy(p, src, bytes, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
                return p;
            };
            D.fsi_pelem = static_cast<const uint32_t*>(put(pelem.data(), pelem.size() * 4));
            D.fsi_unit = static_cast<const double*>(put(P.unit.data(), P.unit.size() * 8));
            D.fsi_econn = static_cast<const uint32_t*>(put(econn.data(), econn.size() * 4));
            D.fsi_eptr = static_cast<const uint32_t*>(put(P.eptr.data(), P.eptr.size() * 4));
            D.fsi_nptr = static_cast<const uint32_t*>(put(nptr.data(), nptr.size() * 4));
            D.fsi_nent = static_cast<const uint32_t*>(put(nent.data(), nent.size() * 4));
            D.fsi_node_of = static_cast<const int32_t*>(put(node_of.data(), node_of.size() * 4));
            double* force = static_cast<double*>(J.arena.alloc((size_t)np * 24));
            D.fsi_nforce = static_cast<double*>(J.arena.alloc((size_t)nn * 24));
            D.fsi_coord = static_cast<double*>(J.arena.alloc((size_t)np * 24));
            D.fsi_vel = static_cast<double*>(J.arena.alloc((size_t)np * 24));
            if (!D.fsi_pelem || !D.fsi_unit || !D.fsi_econn || !D.fsi_eptr || !D.fsi_nptr || !D.fsi_nent ||
                !D.fsi_node_of || !force || !D.fsi_nforce || !D.fsi_coord || !D.fsi_vel)
                return fail(ENSC_ERR_MEM, "ensc_engine_fsi_setup: allocation failed");
            // the particle forces start at zero (resizeFsiSharedData, 1540-1567)
            HIPCHK(hipMemset(force, 0, (size_t)np * 24));
            D.fsi_force = force;
            D.n_fsi_part = np;
            D.n_fsi_node = nn;
        }
        E->dirty = true;
        if ((st = relayout(E))) return st;
        E->fsi_np = np;
        E->fsi_nn = nn;
        if (np && !E->live.empty()) {
            // the particles' positions and velocities now (generation,
            // generateVirtualParticlesForFace 1495-1520)
            hipLaunchKernelGGL(k_fsi_particles, dim3((np + FSI_BLK - 1) / FSI_BLK), dim3(FSI_BLK), 0, E->stream,
                               static_cast<const JobDev*>(E->jobs_p));
            HIPCHK(hipGetLastError());
        }
        if ((st = sync(E))) return st;
        if (n_particles) *n_particles = np;
        return ENSC_OK;
    });
}

ensc_status ensc_engine_fsi_set_forces(ensc_engine* E, const double* forces, int64_t n) {
    if (!E || (!forces && n)) return fail(ENSC_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(E->device));
    Job* J = E->slots.size() == 1 ? job_at(E, 0) : nullptr;
    if (!J || n != (int64_t)J->dev.n_fsi_part)
        return fail(ENSC_ERR_ARG, "ensc_engine_fsi_set_forces: n must equal the particle count of ensc_engine_fsi_setup");
    if (!n) return ENSC_OK;
    // ordered after the queued steps on the engine's stream; returns once copied
    HIPCHK(hipMemcpyAsync(const_cast<double*>(J->dev.fsi_force), forces, (size_t)n * 24, hipMemcpyHostToDevice,
                          E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
    return ENSC_OK;
}

ensc_status ensc_engine_fsi_get(ensc_engine* E, double* coords, double* vel, int64_t n) {
    if (!E) return fail(ENSC_ERR_ARG, "null engine");
    HIPCHK(hipSetDevice(E->device));
    Job* J = E->slots.size() == 1 ? job_at(E, 0) : nullptr;
    if (!J || n != (int64_t)J->dev.n_fsi_part)
        return fail(ENSC_ERR_ARG, "ensc_engine_fsi_get: n must equal the particle count of ensc_engine_fsi_setup");
    ensc_status st = sync(E);
    if (st) return st;
    if (n && coords) HIPCHK(hipMemcpy(coords, J->dev.fsi_coord, (size_t)n * 24, hipMemcpyDeviceToHost));
    if (n && vel) HIPCHK(hipMemcpy(vel, J->dev.fsi_vel, (size_t)n * 24, hipMemcpyDeviceToHost));
    return ENSC_OK;
}

ensc_status ensc_engine_fsi_device(ensc_engine* E, double** forces, const double** coords, const double** vel,
                                   void** stream) {
    if (!E) return fail(ENSC_ERR_ARG, "null engine");
    Job* J = E->slots.size() == 1 ? job_at(E, 0) : nullptr;
    if (!J || !J->dev.n_fsi_part) return fail(ENSC_ERR_STATE, "ensc_engine_fsi_device: no particles (ensc_engine_fsi_setup)");
    if (forces) *forces = const_cast<double*>(J->dev.fsi_force);
    if (coords) *coords = J->dev.fsi_coord;
    if (vel) *vel = J->dev.fsi_vel;
    if (stream) *stream = E->stream;
    return ENSC_OK;
}

int64_t ensc_engine_device_bytes(ensc_engine* E) {
    if (!E) return 0;
    int64_t t = 0;
    for (auto& j : E->slots)
        if (j) t += (int64_t)j->arena.total;
    return t + (int64_t)(E->jobs_buf.cap + E->b2e_buf.cap + E->b2n_buf.cap + E->b2f_buf.cap + E->pe_buf.cap +
                         E->pn_buf.cap + E->b2fin_buf.cap + E->fin_buf.cap) +
           (int64_t)E->ctl_cap * (int64_t)sizeof(Ctl);
}

uint32_t ensc_batch_size(ensc_engine* E) { return E ? (uint32_t)E->slots.size() : 0; }

ensc_status ensc_batch_job_scalars(ensc_engine* E, uint32_t job, double* out) {
    if (!E || !out) return fail(ENSC_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(E->device));
    return engine_job_scalars(E, job, out);
}

ensc_status ensc_batch_job_download(ensc_engine* E, uint32_t job, const char* field, double* out, int64_t count) {
    if (!E || !field || !out) return fail(ENSC_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(E->device));
    return engine_job_download(E, job, field, out, count);
}

}  // extern "C"
