/*
 * oracle.h - C API of the CPU FP64 restatement of the reference step loop.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg may load liboracle.so, and only as the checker / the
 * timed CPU baseline.  The engine never links, loads or falls back to it.
 */
#ifndef ENSC_ORACLE_H
#define ENSC_ORACLE_H

#include <stdint.h>
#include "../include/ensc_model.h"

#ifdef __cplusplus
extern "C" {
#endif

/* History row layout written by oracle_run: ORACLE_HIST_BASE scalars, then
 * 3 displacement components per probe node. */
enum {
    OH_TIME = 0, OH_DT_I, OH_DT_I_1, OH_DTS_MIN, OH_KE, OH_IE, OH_HG, OH_VIS, OH_WEXT,
    ORACLE_HIST_BASE
};

/* Scalars returned by oracle_scalars. */
enum {
    OS_TIME = 0, OS_DT_I, OS_DT_I_1, OS_DTS_MIN, OS_KE, OS_IE, OS_HG, OS_VIS, OS_WEXT,
    OS_KE0, OS_NSTEPS, OS_STEP_INDEX, OS_LMIN, OS_COUNT
};

void* oracle_create(const ensc_model_desc* m, int n_threads);
void oracle_destroy(void* h);
/* exDyna3D::init (exDyna3D.cpp:105-142). */
int oracle_init(void* h);
/* exDyna3D::run's step loop (exDyna3D.cpp:76-101), stopping after
 * max_steps computeSate calls when max_steps > 0.  Writes one history row
 * per step into hist (rows of ORACLE_HIST_BASE + 3*n_probe doubles) while
 * hist_rows allows.  Returns the number of steps taken. */
int64_t oracle_run(void* h, int64_t max_steps, const uint32_t* probe, uint32_t n_probe,
                   double* hist, int64_t hist_rows);
/* Enter the next *Step (setCurrentStep and its end time) without stepping,
 * as oracle_run does before its first step there. */
void oracle_begin_step(void* h);
/* Time n_steps iterations of the hot loop body (dt update + computeSate)
 * inside the current *Step, with a steady clock around the loop only.
 * Returns seconds.  Used as the CPU baseline. */
double oracle_time_steps(void* h, int64_t n_steps);
/* Copy a field out: "u","v","x","a","rf" ([Nn][3]); "sigma" ([Ne][6]);
 * "qia" ([Ne][12]); "epsp","dtStable","Length","V0" ([Ne]);
 * "B0" ([Ne][8][3]); "invM" ([Nn]).  Returns count copied or -1. */
int64_t oracle_get(void* h, const char* field, double* out);
void oracle_scalars(void* h, double* out /* OS_COUNT */);
int oracle_threads(void* h);
/* output_results (exDyna3D.cpp:207-336): record a snapshot at every output
 * time (interval <= 0: calculate_time_interval's rule, 1947-1977). */
void oracle_set_output(void* h, double interval);
int64_t oracle_snapshots(void* h);
/* Fluid-structure coupling (exDyna3D.cpp:1281-1758): virtual particles on
 * the current state, n_layers per FSI face, dist apart (returns the count);
 * their forces [n][3], applied at every later step; their positions and
 * velocities [n][3] (returns n). */
int64_t oracle_fsi_setup(void* h, int n_layers, double dist);
int oracle_fsi_set_forces(void* h, const double* forces, int64_t n);
int64_t oracle_fsi_get(void* h, double* coords, double* vel);
/* Field of snapshot k: "time" [1]; "x","a","v","u" [Nn][3]; "vonMises" [Ne];
 * "Stress", "Green-Lagrange Strain" [Ne][6] (xx yy zz xy yz xz).  Returns
 * the count written or -1. */
int64_t oracle_snapshot(void* h, int64_t k, const char* field, double* out);

#ifdef __cplusplus
}
#endif
#endif
