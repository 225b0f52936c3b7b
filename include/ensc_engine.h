/*
 * ensc_engine.h - C ABI of the MI355X explicit-dynamics step engine.
 *
 * Drop-in for the CPU step loop of ssadz/EnSCDynamics' exDyna3D
 * (Explicit3D/include/exDyna3D.h:36-222).  The reference exposes a C++
 * class with public run()/init()/setCurrentStep() and protected virtual
 * computeSate()/compute_a() hooks; a binding that replaces its step loop
 * calls these entry points instead.
 *
 * One ensc_engine holds one or more jobs (meshes) resident in device
 * memory.  All calls are host-thread-safe per engine only when serialized by
 * the caller.  Calls that touch the device make the engine's device
 * (ensc_options.device) current on the calling thread and leave it current,
 * so engines on different GPUs can be driven from one thread.  Every call
 * returns an ensc_status; ensc_last_error() gives the message of the last
 * failure on the calling thread.
 */
#ifndef ENSC_ENGINE_H
#define ENSC_ENGINE_H

#include <stdint.h>
#include "ensc_model.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ensc_status {
    ENSC_OK = 0,
    ENSC_ERR_ARG = 1,      /* bad argument or model */
    ENSC_ERR_HIP = 2,      /* HIP runtime error */
    ENSC_ERR_STATE = 3,    /* call out of sequence */
    ENSC_ERR_MEM = 4,      /* device memory exhausted / over budget */
    ENSC_ERR_NUMERIC = 5,  /* a job raised a device error flag */
    ENSC_ERR_NODEV = 6     /* no GPU or the kernels are missing */
} ensc_status;

typedef enum ensc_precision {
    ENSC_FP64 = 0,   /* everything FP64: parity path (<=1e-10 vs the reference) */
    ENSC_MIXED = 1   /* FP32 element kinematics/constitutive, FP64 nodal state,
                        energies and dt; compensated position accumulation.
                        Quasi-static jobs run in FP64 (ensc_options.fp64_jobs) */
} ensc_precision;

/* Device error flags of a job (ensc_engine_scalars ENSC_S_ERR). */
enum {
    ENSC_EF_JACOBIAN = 1u,  /* det F <= 0 in some element */
    ENSC_EF_NONFINITE = 2u, /* non-finite element force or stress */
    ENSC_EF_DT = 4u         /* stable dt collapsed (<= 0 or non-finite) */
};

typedef struct ensc_options {
    int32_t precision;      /* ensc_precision */
    int32_t reorder;        /* 1: renumber elements and nodes for locality
                               (space-filling order); results are returned in
                               the caller's numbering either way */
    int32_t block_elem;     /* element-kernel block size (0 = autotuned/default) */
    int32_t block_node;     /* node-kernel block size (0 = default) */
    int32_t device;         /* HIP device ordinal */
    uint32_t n_probe;       /* nodes whose displacement is recorded per step */
    const uint32_t* probe_nodes;
    int64_t hist_cap;       /* rows of per-step history kept on device */
    int32_t staging;        /* 0: stage element-block nodes in LDS when they
                               fit; mixed jobs with block assembly and at
                               most 512 nodes per element block run the
                               persistent pipelined element pass (default);
                               1: staged, one block per element block;
                               2: gather corners from HBM */
    int32_t waves;          /* element-kernel occupancy target, waves per SIMD
                               (1, 2 or 3; 0 = default for the precision).
                               Staged kernels only. */
    int32_t fp64_jobs;      /* mixed engine, per-job precision rule (decided
                               from each job alone, so a job computes the same
                               alone and in any mix):
                               0: quasi-static jobs - runs spanning >= 512
                                  wave transits (total time x wave speed /
                                  largest extent of the mesh) - run the FP64
                                  kernels, every other job mixed precision;
                               1: every job mixed precision;
                               2: every job FP64.
                               FP64 engines run every job in FP64. */
    int32_t fp64_order;     /* the order of FP64 jobs' arithmetic:
                               0: the element body regrouped through the
                               corners' Walsh sign patterns, and each element
                               block sums its elements' corner forces in LDS,
                               the node adding the blocks' partials (no
                               per-corner forces through HBM; <= 1e-10 of the
                               reference, rounding of the other order);
                               1: the reference's order, every expression of
                               the element body as exDyna3D.cpp:838-1109
                               writes it and the node's sum per corner then
                               ascending element (584-639) through the gather
                               table - bit identical to the CPU restatement. */
} ensc_options;

/* Scalars of one job (ensc_engine_scalars). */
enum {
    ENSC_S_TIME = 0, ENSC_S_DT_I, ENSC_S_DT_I_1, ENSC_S_DTS_MIN, ENSC_S_KE, ENSC_S_IE, ENSC_S_HG,
    ENSC_S_VIS, ENSC_S_WEXT, ENSC_S_KE0, ENSC_S_NSTEPS, ENSC_S_STEP_INDEX, ENSC_S_STEP_END,
    ENSC_S_ERR, ENSC_S_DONE, ENSC_S_COUNT
};

/* One history row: ENSC_H_BASE scalars then 3 displacements per probe. */
enum {
    ENSC_H_TIME = 0, ENSC_H_DT_I, ENSC_H_DT_I_1, ENSC_H_DTS_MIN, ENSC_H_KE, ENSC_H_IE, ENSC_H_HG,
    ENSC_H_VIS, ENSC_H_WEXT, ENSC_H_BASE
};

typedef struct ensc_engine ensc_engine;

const char* ensc_last_error(void);
/* Number of visible HIP devices (0 when none). */
int32_t ensc_device_count(void);
/* Measurement utility (no counterpart in the reference): times `reps`
 * launches of a plain device-to-device copy of `bytes` bytes on `device`
 * and returns the read+write bandwidth in GB/s.  Its kernel, k_copy_probe,
 * is the known-size reference for scaling hardware byte counters. */
ensc_status ensc_copy_bandwidth(int32_t device, int64_t bytes, int32_t reps, double* gbs);

/* Lifecycle.  create() copies the model; the caller keeps its arrays.
 * Replaces the reference's DataIn population + init_data + compute_invMassMatrix
 * (exDyna3D.cpp:144-205, 492-515). */
ensc_status ensc_engine_create(const ensc_model_desc* model, const ensc_options* opts, ensc_engine** out);
ensc_status ensc_engine_destroy(ensc_engine* eng);

/* exDyna3D::init (exDyna3D.cpp:105-142): initial velocity and the
 * zero-increment element pass that fills the stable-dt array. */
ensc_status ensc_engine_init(ensc_engine* eng);

/* exDyna3D::setCurrentStep (exDyna3D.cpp:1760-1940) for step_index >= 1,
 * then stepEndTime = time + timePeriod (exDyna3D.cpp:83). */
ensc_status ensc_engine_set_step(ensc_engine* eng, uint32_t step_index);

/* Run the `while (time < stepEndTime)` loop of exDyna3D::run
 * (exDyna3D.cpp:86-100) for at most max_steps iterations (<= 0: until the
 * step ends).  *done is set to 1 when time has reached stepEndTime. */
ensc_status ensc_engine_advance(ensc_engine* eng, int64_t max_steps, int32_t* done);

/* ensc_engine_advance for callers that cannot allocate device memory (code
 * run inside stream capture, tight serving loops, hosts that forbid
 * hipMalloc).  No counterpart in the reference.  Same loop, same results
 * bit for bit; the call allocates nothing: the job table, block maps and
 * block partials of its step iterations live in a static 4 MiB device array
 * (one per GPU, held until the process ends; every call on a GPU shares it
 * and rewrites it first, so calls that overlap on one GPU race), and it
 * launches plainly instead of instantiating a hipGraph.  Fails with
 * ENSC_ERR_ARG when the engine's tables need more than the array (about 10M
 * elements in one mesh) and with ENSC_ERR_STATE when output snapshots are
 * on.  Later ensc_engine_advance calls use the engine's own buffers again. */
ensc_status ensc_engine_advance_static(ensc_engine* eng, int64_t max_steps, int32_t* done);

/* exDyna3D::run (exDyna3D.cpp:54-102): init, then every *Step. */
ensc_status ensc_engine_run(ensc_engine* eng);

/* Enqueue n step iterations on the engine's stream without waiting
 * (iterations after the step end are no-ops on the device). */
ensc_status ensc_engine_enqueue_steps(ensc_engine* eng, int64_t n);
ensc_status ensc_engine_sync(ensc_engine* eng);

/* Copy a field to host memory, in the caller's node/element numbering:
 * "u","v","x","rf" -> [n_nodes][3]; "sigma" -> [n_elems][6] (xx yy zz xy xz yz);
 * "qia" -> [n_elems][12]; "epsp" -> [n_elems].
 * count is the number of doubles the caller's buffer holds. */
ensc_status ensc_engine_download(ensc_engine* eng, const char* field, double* out, int64_t count);
ensc_status ensc_engine_scalars(ensc_engine* eng, double* out /* ENSC_S_COUNT */);
/* Per-step history rows recorded so far (up to rows). */
ensc_status ensc_engine_history(ensc_engine* eng, double* out, int64_t rows, int64_t* got);

/* Element/node kernel timing over the last enqueue_steps call, measured with
 * HIP events on the engine's stream: mean ms per launch of each kernel and
 * the number of launches timed. */
ensc_status ensc_engine_kernel_times(ensc_engine* eng, double* elem_ms, double* node_ms, int64_t* launches);
/* Mean duration in ms of each launch of the last timed enqueue, in launch
 * order: element pass, element finalize, nodal update, node finalize. */
ensc_status ensc_engine_phase_times(ensc_engine* eng, double* ms4);
/* Launch layout of the last step iteration (ENSC_L_COUNT values): whether
 * element blocks stage their nodes in LDS, the element blocks run by the
 * pipelined pass and by the one-block-per-tile pass, the node blocks, the
 * dynamic LDS bytes per element block of the mixed jobs, the jobs using
 * block assembly and the jobs running in FP64. */
enum {
    ENSC_L_STAGED = 0, ENSC_L_PIPE_BLOCKS, ENSC_L_ELEM_BLOCKS, ENSC_L_NODE_BLOCKS, ENSC_L_LDS_BYTES,
    ENSC_L_ASM_JOBS, ENSC_L_FP64_JOBS, ENSC_L_COUNT
};
ensc_status ensc_engine_layout(ensc_engine* eng, int64_t* out /* ENSC_L_COUNT */);
/* ---- output snapshots: exDyna3D::output_results (exDyna3D.cpp:207-336) ----
 * Single-job engines.  At every step whose time reaches the next output time
 * (the first step after enabling, then every `interval`) or the total time,
 * the device captures what output_results writes: node positions, a (after
 * compute_a), v (after the velocity BCs) and u, all before the step's
 * velocity update, and the element stress.  Snapshots go to a ring of
 * `slots` device buffers (1..64; 0 turns output off).  interval <= 0 takes
 * the reference's rule (calculate_time_interval, exDyna3D.cpp:1947-1977).
 * With `dir` set, ensc_engine_advance / ensc_engine_run write each snapshot
 * as <dir>/solution_<n>.vtu (n = 0, 1, ...; include/ensc_vtu.h) as they
 * go, like the reference's ResultData/solution_<n>.vtu.  With dir NULL the
 * caller drains the ring (output_pending / output_field / output_write /
 * output_pop); a job waits while every slot is full, and advance returns
 * with *done = 0 so the caller can drain. */
ensc_status ensc_engine_set_output(ensc_engine* eng, double interval, int32_t slots, const char* dir);
/* Undrained snapshots and the time of the oldest (time may be NULL). */
ensc_status ensc_engine_output_pending(ensc_engine* eng, int32_t* n, double* time);
/* A field of the oldest snapshot, caller's numbering: "x", "a", "v", "u"
 * [n_nodes][3]; "vonMises" [n_elems]; "Stress", "Green-Lagrange Strain"
 * [n_elems][6] in the order the reference hands to VTK (xx yy zz xy yz xz).
 * Green-Lagrange strain is 0.5 (F^T F - I) with the step's F = I + u B0. */
ensc_status ensc_engine_output_field(ensc_engine* eng, const char* field, double* out, int64_t count);
/* Write the oldest snapshot as a .vtu file with the reference's arrays. */
ensc_status ensc_engine_output_write(ensc_engine* eng, const char* path);
/* Drop the oldest snapshot. */
ensc_status ensc_engine_output_pop(ensc_engine* eng);

/* Autotuner (extension; no counterpart in the reference).  Times `ticks`
 * step iterations (1..10000, after two untimed ones) of the engine's live
 * jobs under each candidate launch configuration and keeps the fastest:
 * blocks per CU of the pipelined element pass (1 or 2), waves per SIMD of
 * the staged element kernel (1, 2 or 3) and the nodal update's block size
 * (64, 128 or 256), searched one after the other.  None of these changes a
 * job's tables or its element and nodal arithmetic; the node block size
 * regroups only the kinetic-energy and external-work block sums, so those
 * two diagnostics may change in their last bits.  Call it after
 * ensc_engine_init and ensc_engine_set_step: the trial steps run from the
 * current state, which is saved on the device first and restored after
 * every trial, so the jobs continue from where they were.  chosen (may be
 * null) receives {waves, node block size, pipelined blocks per CU, ns per
 * step iteration of the choice}. */
ensc_status ensc_engine_autotune(ensc_engine* eng, int32_t ticks, int64_t* chosen);


/* ---- fluid-structure coupling through virtual particles ----
 * The reference's FSI interface (exDyna3D.cpp:1281-1758): the public
 * fsi_share_data member (exDyna3D.h:54, FSI_share_data.h:27-36) holds one
 * position, velocity and force per virtual particle; a fluid code writes
 * the forces and reads the particles between steps.  Single-job engines.
 *
 * fsi_setup replaces get_fsiSph_virtualParticles_and_vel(nLayers, dist)
 * (1363-1383; run() calls it with 0 layers, 67): boundary nodes are those
 * held by fewer than 8 elements, FSI faces those with 4 boundary corners
 * (1282-1361), and each FSI face gets n_layers particles stepping inward
 * from its centre, dist apart (1428-1521), on the current positions;
 * particles are ordered by element, face, layer as in the reference.  Call
 * after ensc_engine_init; calling again replaces the particles.  The
 * particle forces start at zero and persist until set again.  Each step
 * adds them to the nodes through the shape functions before the internal
 * forces (apply_fsiSph_nodeForce, 1635-1689, called at 783) and moves the
 * particles with the mesh after the nodal update
 * (update_virParticles_coor_vel, 1691-1758, called at 792); their work
 * enters the external work. */
ensc_status ensc_engine_fsi_setup(ensc_engine* eng, int32_t n_layers, double dist, int64_t* n_particles);
/* fsi_share_data.FSI_virtualParticles_nodeForce <- forces [n][3] (host),
 * n = the particle count; returns once copied. */
ensc_status ensc_engine_fsi_set_forces(ensc_engine* eng, const double* forces, int64_t n);
/* FSI_virtualParticles_coordinates / _velocity -> coords, vel [n][3] (host;
 * either may be NULL), after the queued steps. */
ensc_status ensc_engine_fsi_get(ensc_engine* eng, double* coords, double* vel, int64_t n);
/* Device pointers of the same arrays ([n][3] FP64) and the engine's stream
 * (a hipStream_t), for a coupler that runs on the GPU: work it queues on
 * that stream between steps is ordered with the engine's (no host copies). */
ensc_status ensc_engine_fsi_device(ensc_engine* eng, double** forces, const double** coords, const double** vel,
                                   void** stream);

/* Enable (1) / disable (0) per-kernel event timing for later enqueue calls. */
ensc_status ensc_engine_set_timing(ensc_engine* eng, int32_t on);
/* Device bytes held by the engine. */
int64_t ensc_engine_device_bytes(ensc_engine* eng);
/* Capture `ticks` step iterations (the four launches of each) as a
 * hipGraph and replay it for enqueued steps (0: plain launches).  The graph
 * is re-captured whenever the set of stepped jobs changes. */
ensc_status ensc_engine_set_graph(ensc_engine* eng, int32_t ticks);

/* ---- batches: many independent meshes in one engine ----
 * No counterpart in the reference (it runs one mesh per process).  Every job
 * keeps its own time step, step end and BCs; one launch of each kernel per
 * step iteration covers all unfinished jobs, whose element and node blocks
 * are packed back to back.  ensc_engine_init / set_step / advance / run apply
 * to every job of the engine; jobs are numbered 0..n-1 in creation order. */
ensc_status ensc_batch_create(const ensc_model_desc* const* models, uint32_t n_models, const ensc_options* opts,
                              ensc_engine** out);
uint32_t ensc_batch_size(ensc_engine* eng);
ensc_status ensc_batch_job_scalars(ensc_engine* eng, uint32_t job, double* out /* ENSC_S_COUNT */);
ensc_status ensc_batch_job_download(ensc_engine* eng, uint32_t job, const char* field, double* out, int64_t count);

/* ---- continuous-batching solve server ----
 * No counterpart in the reference.  Requests queue on the host and worker
 * threads prepare them (setup arithmetic, renumbering, gather and staging
 * tables) while the GPU steps.  Each pump admits prepared jobs in arrival
 * order while their device footprint fits hbm_budget_bytes (and at most
 * max_live_jobs are live), runs their init pass inside the next step
 * iteration of the live batch, advances every live job together for
 * `ticks` step iterations, switches a job that reached the end of one
 * *Step to the next (setCurrentStep, as exDyna3D::run does), and retires
 * jobs that reached the end of their last step or raised an error flag
 * (results to the host at once; their device memory is freed in groups,
 * when retired jobs hold a quarter of the launch, when nothing is live or
 * when an admission needs the room).  With nothing on the device,
 * a pump waits for the next job's preparation.  A job whose set-up fails
 * (e.g. an inverted element) is retired at once and ensc_server_result
 * returns ENSC_ERR_ARG with the reason.
 * Times are seconds on the server's steady clock (ensc_server_clock). */
typedef struct ensc_server ensc_server;
ensc_status ensc_server_create(const ensc_options* opts, int64_t hbm_budget_bytes, int32_t max_live_jobs,
                               ensc_server** out);
ensc_status ensc_server_destroy(ensc_server* srv);
ensc_status ensc_server_set_graph(ensc_server* srv, int32_t ticks);
/* Copy a job into the queue.  keep_result != 0 keeps its final displacement
 * field for ensc_server_result.  t_arrival < 0 means now. */
ensc_status ensc_server_submit(ensc_server* srv, const ensc_model_desc* model, int32_t keep_result, double t_arrival,
                               uint64_t* ticket);
/* Admit, step, retire.  Writes up to cap retired tickets to done. */
ensc_status ensc_server_pump(ensc_server* srv, int32_t ticks, uint64_t* done, uint32_t cap, uint32_t* n_done);
ensc_status ensc_server_result(ensc_server* srv, uint64_t ticket, double* scalars /* ENSC_S_COUNT */, double* u,
                               int64_t u_count, double* times /* submit, admit, finish */);
/* bytes_used: device bytes of live jobs and of retired jobs not yet freed. */
ensc_status ensc_server_status(ensc_server* srv, uint32_t* queued, uint32_t* live, int64_t* bytes_used);
double ensc_server_clock(ensc_server* srv);

#ifdef __cplusplus
}
#endif

#endif /* ENSC_ENGINE_H */
