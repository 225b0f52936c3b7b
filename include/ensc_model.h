/*
 * ensc_model.h - plain-C description of one explicit-dynamics job.
 *
 * This is the data that the reference's DataIn fills into exDyna3D's
 * protected members (Explicit3D/include/exDyna3D.h:141-220, written by
 * Explicit3D/source/dataIn.cpp:1550-1788), flattened into arrays so it can
 * cross a C ABI.  Node sets are already resolved to node ids: the reference
 * looks a set name up on every BC application (exDyna3D.cpp:412-419,
 * 464-470), but sets never change after reading, so resolving once is the
 * same.  An unknown set name resolves to an empty list (the reference warns
 * and skips it, exDyna3D.cpp:417).
 *
 * All indices are 0-based.  Arrays are borrowed: the consumer copies what it
 * keeps.
 */
#ifndef ENSC_MODEL_H
#define ENSC_MODEL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One boundary-condition line of a *Boundary block
 * (BoundaryCondition.h:13-19: spc_nodes / vel_nodes entries). */
typedef struct ensc_bc_entry {
    const uint32_t* nodes; /* node ids the line names (set or single node) */
    uint32_t n_nodes;
    int32_t d0, d1;  /* inclusive dof range, 0..2 */
    int32_t cond;    /* SPC: third numeric field; only cond==1 constrains
                        (exDyna3D.cpp:478, 2129).  VEL: ignored. */
    double value;    /* VEL: prescribed velocity (exDyna3D.cpp:428) */
    int32_t amp;     /* VEL: amplitude index, -1 for none (extension: the
                        reference keeps velocity BCs constant) */
    int32_t pad_;
} ensc_bc_entry;

/* One *Dsload line with its surface resolved to element faces (extension:
 * the reference parses *Dsload and *Surface, dataIn.cpp:1339-1478, and never
 * applies them).  A pressure `value` (times the amplitude) acts on each face
 * over its current area, into the element for a positive value. */
typedef struct ensc_dsload_entry {
    const uint32_t* elems;  /* element of each loaded face */
    const int32_t* faces;   /* its face: 0..5 = S1..S6 (C3D8 face numbering) */
    uint32_t n_faces;
    int32_t amp;            /* amplitude index, -1 for none */
    double value;           /* pressure */
} ensc_dsload_entry;

/* One *Step (StepData.h:14-34).  steps[0] is the reference's "Initial"
 * step holding BCs defined outside any *Step (dataIn.cpp:323-326). */
typedef struct ensc_step_def {
    double period;               /* timePeriod from *Dynamic, Explicit */
    int32_t reset_spc;           /* *Boundary, op=NEW */
    int32_t reset_vel;           /* *Boundary, op=NEW, type=VELOCITY */
    int32_t reset_dload;         /* *Dload, op=NEW */
    int32_t grav_on;             /* *Dload ... GRAV */
    double grav_value;
    double grav_dir[3];
    int32_t grav_amp;            /* amplitude index or -1 */
    uint32_t n_spc;
    const ensc_bc_entry* spc;
    uint32_t n_vel;
    int32_t pad_;
    const ensc_bc_entry* vel;
    /* surface pressures (extension); inherited like the reference's dsload
     * list (exDyna3D.cpp:1897-1907): a step without lines keeps the previous
     * step's unless reset_dsload (*Dsload, op=NEW) */
    int32_t reset_dsload;
    uint32_t n_dsload;
    const ensc_dsload_entry* dsload;
} ensc_step_def;

/* *Amplitude table: pairs (t_i, a_i) (dataIn.cpp:1242-1271). */
typedef struct ensc_amplitude {
    const double* t;
    const double* a;
    uint32_t n;
    int32_t pad_;
} ensc_amplitude;

typedef struct ensc_model_desc {
    uint32_t n_nodes;
    uint32_t n_elems;
    const double* x0;      /* [n_nodes][3] initial coordinates */
    const uint32_t* conn;  /* [n_elems][8] C3D8R corners, Element_HexN8.cpp:50-89 order */
    double E, nu, rho;     /* Material_elastic (Material_elastic.h:8-21) */
    int32_t plastic;       /* extension: J2 with linear isotropic hardening */
    int32_t pad0_;
    double yield0, hard;   /* initial yield stress, plastic modulus */
    double c_dr, c_cr;     /* hourglass stiffness / viscosity (exDyna3D.h:182-183) */
    double factor;         /* factor_timeStep */
    double cvisl, cvisq;   /* *Bulk Viscosity */
    double total_time;     /* sum of step periods (dataIn.cpp:869) */
    const double* v_init;  /* [n_nodes][3] velocity after *Initial Conditions */
    uint32_t n_steps;
    uint32_t n_amps;
    const ensc_step_def* steps;
    const ensc_amplitude* amps;
} ensc_model_desc;

/* Material constants exactly as Material_elastic::update computes them
 * (Material_elastic.cpp:9-14). */
typedef struct ensc_material_consts {
    double G, K, lambda, wos;
} ensc_material_consts;

#ifdef __cplusplus
}
#endif

#endif /* ENSC_MODEL_H */
