// Device-side data structures shared by the engine's kernels and host code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ensc {

// Typed address spaces for device code.  Pointers read from the job table
// are generic; casting them to the global (1) / LDS (3) space lets the
// compiler emit global_load/ds_read with an SGPR base and a 32-bit VGPR
// offset instead of flat accesses with 64-bit VGPR addresses.  Byte offsets
// are formed in 32 bits: every per-job array is kept under 4 GiB (checked at
// job build).
#define ENSC_GLB __attribute__((address_space(1)))
#define ENSC_LDS __attribute__((address_space(3)))
template <class T>
__device__ __forceinline__ const ENSC_GLB T& gld(const T* base, uint32_t idx) {
    return *(const ENSC_GLB T*)((const ENSC_GLB char*)base + (uint32_t)(idx * (uint32_t)sizeof(T)));
}
template <class T>
__device__ __forceinline__ ENSC_GLB T& gst(T* base, uint32_t idx) {
    return *(ENSC_GLB T*)((ENSC_GLB char*)base + (uint32_t)(idx * (uint32_t)sizeof(T)));
}
// 16-byte pair of doubles: node records are read and written two values per
// access (records are 80 bytes, so every even index is 16-byte aligned).
typedef double dbl2 __attribute__((ext_vector_type(2)));
struct D2 {
    double a, b;
};
__device__ __forceinline__ D2 gld2(const double* base, uint32_t idx) {
    const dbl2 v = *(const ENSC_GLB dbl2*)((const ENSC_GLB char*)base + (uint32_t)(idx * 8u));
    return D2{v.x, v.y};
}
__device__ __forceinline__ void gst2(double* base, uint32_t idx, double a, double b) {
    dbl2 v;
    v.x = a;
    v.y = b;
    *(ENSC_GLB dbl2*)((ENSC_GLB char*)base + (uint32_t)(idx * 8u)) = v;
}

// What the element kernel did in the current step iteration ("tick").
enum : int32_t { KIND_NONE = 0, KIND_STEP = 1, KIND_PREPASS = 2, KIND_INIT = 3 };

// BC mask bits per node: bit d = velocity BC on dof d, bit 3+d = SPC on dof d.
constexpr uint8_t BC_VEL0 = 1u, BC_SPC0 = 8u;
constexpr uint32_t ELL_EMPTY = 0xFFFFFFFFu;
// Node state, two arrays of 16-byte-aligned records:
//   pos [nn][4] = [X0 X1 X2 invM]   (32 bytes)
//   vu  [nn][6] = [v0 v1 v2 u0 u1 u2] (48 bytes)
// X is the current position in the FP64 engine (accumulated like the
// reference's vertices) and the reference position in the mixed engine,
// where the current position is X + u and pos is never rewritten.  Keeping
// v and u apart from X makes the nodal update's stores contiguous.
constexpr int POS_REC = 4, VU_REC = 6;
constexpr int NODE_BYTES = 8 * (POS_REC + VU_REC);
constexpr int MAX_PROBE = 8;

// Per-job scalar state, resident on the device.  The time-step scalars are
// the reference's (exDyna3D.h:175-179); dts_min is dtStable.minCoeff() of
// the last element pass (exDyna3D.cpp:677).
struct Ctl {
    double time, dt_i, dt_i_1, step_end, total_time, dts_min, lmin_pre;
    double ke, ie, hg, vis, wext, ke0;
    double grav_amp;
    long long nsteps;
    long long step_limit;  // no step starts once nsteps reaches it
    int32_t kind, have_lpre, init_pending, step_index;
    uint32_t err, pad;
    // output snapshots (exDyna3D::output_results, exDyna3D.cpp:207-219):
    // next output time and interval; a ring of snap_cap slots of which
    // snap_head have been filled and snap_tail drained by the host; snap_now
    // is 1 + the slot this tick's nodal update fills (0: none)
    double t_out, t_int;
    int32_t out_on, snap_now, snap_head, snap_tail, snap_cap, pad2;
    // what the tick's passes did, for the element finalize's blocks that
    // must not read the fields its deciding block rewrites: tick_nsteps is
    // nsteps as the last nodal update saw it when that update ran a step
    // (its history row is tick_nsteps - 1), 0 when none is open; tick_elem
    // is the kind of the tick's element pass
    long long tick_nsteps;
    int32_t tick_elem, pad3;
};

struct JobDev {
    // element arrays, SoA with stride ne_s (precision-dependent element type)
    const uint32_t* conn;  // [8][ne_s]
    const void* Gi;        // [9][ne_s]  Gi(i,j) = (dX/dxi)^-1(i,j) / 8 at row 3i+j;
                           //   B0(k,i) = sum_j Gi(i,j) * sign(k,j) (FEValues at the centre)
    const void* V0;        // [ne_s]
    void* sigma;           // [6][ne_s]
    void* qia;             // [12][ne_s]
    void* epsp;            // [ne_s]
    void* fel;             // [8][ne_s][3] element nodal forces (corner-major, packed xyz);
                           //   gathered assembly only
    // LDS staging of each element block's nodes (null when not staged)
    const uint16_t* lconn;     // [4][ne_s][2] corner -> index in the block's node list, corners
                               //   2p and 2p + 1 of an element side by side (one 32-bit word)
    const uint32_t* blk_ptr;   // [nblk_e + 1] offsets into blk_nodes
    const uint32_t* blk_nodes; // node ids of each element block, ascending
    // block assembly: each element block sums its elements' corner forces
    // per block node in LDS and writes one partial per (block, node) at the
    // node's position in blk_nodes; the nodal update adds a node's partials
    // in ascending position order.  FP32 partials for mixed jobs, FP64 for
    // FP64 jobs.
    void* part_f;              // [n_pos][3] partial nodal forces
    const uint32_t* pos_ptr;   // [nn + 1] offsets into pos_list
    const uint32_t* pos_list;  // positions in blk_nodes of each node, ascending
    // node arrays
    double* pos;           // [nn][POS_REC]  X, invM
    double* vu;            // [nn][VU_REC]   v, u
    const double* inv_m;   // [nn] 1/m (mixed; the nodal update reads it without the position record)
    uint32_t* ucomp;       // [nn] compensation of the u sum (mixed; 3 x 10-bit fields) or null
    double* rf;            // [nn][3]
    const uint8_t* bcmask; // [nn]
    const double* bcval;   // [nn][3]
    const int16_t* bcamp;  // [nn][3]
    const uint32_t* ell;   // [ell_w][nn_s] slots (e << 3 | corner), sorted by (corner, original element)
    // amplitude tables, packed: table i is amp_t/amp_a[amp_off[i] .. amp_off[i+1])
    const double* amp_t;
    const double* amp_a;
    const int32_t* amp_off;
    int32_t n_amp;
    // surface pressures of the current step (extension; *Dsload): per loaded
    // face its 4 nodes (inward-normal order), pressure and amplitude; the
    // face-force kernel writes each face's per-node force, and the nodal
    // update adds a node's face forces in ascending face order
    uint32_t n_dsf, pad_dsf;
    const uint32_t* dsf_nodes;  // [n_dsf][4]
    const double* dsf_val;      // [n_dsf]
    const int32_t* dsf_amp;     // [n_dsf]
    double* dsf_force;          // [n_dsf][3]
    const uint32_t* ds_ptr;     // [nn + 1] offsets into ds_list
    const uint32_t* ds_list;    // faces of each node, ascending
    // fluid-structure coupling through virtual particles (exDyna3D.cpp:
    // 1281-1758; single-job engines): each particle sits at fixed unit
    // coordinates of its FSI element; per FSI node the (element, corner)
    // entries in ascending original element order
    uint32_t n_fsi_part, n_fsi_node;
    const uint32_t* fsi_pelem;   // [n_fsi_part] FSI element slot of each particle
    const double* fsi_unit;      // [n_fsi_part][3] unit coordinates
    const uint32_t* fsi_econn;   // [slots][8] node ids of each FSI element
    const uint32_t* fsi_eptr;    // [slots + 1] particle range of each FSI element
    const uint32_t* fsi_nptr;    // [n_fsi_node + 1] offsets into fsi_nent
    const uint32_t* fsi_nent;    // slot << 3 | corner
    const int32_t* fsi_node_of;  // [nn] FSI node index, -1 for none
    const double* fsi_force;     // [n_fsi_part][3] particle forces (set by the coupler)
    double* fsi_nforce;          // [n_fsi_node][3] nodal force of the particles
    double* fsi_coord;           // [n_fsi_part][3] particle positions
    double* fsi_vel;             // [n_fsi_part][3] particle velocities
    // gravity of the current step (StepData.h:22)
    int32_t grav_on, grav_amp;
    double grav_g, grav_dir[3];
    // material and parameters (Material_elastic.cpp:9-14; exDyna3D.h:179-188)
    double lambda, G, K, rho, wos, c_dr, c_cr, cvisl, cvisq, factor, yield0, hard;
    // the same constants folded for the mixed-precision element body
    struct MixConst {
        float lam, two_g, three_g, inv_ret;  // inv_ret = 1 / (3 G + hard)
        float yield0, hard, dr, cr;          // dr = c_dr (4/3 G + K) / 10 / sqrt 8;  cr = c_cr rho wos / 4
        float rho, cq2, clw, pad;            // cq2 = cvisq^2;  clw = cvisl wos
        double inv_wos, cq2_wos;             // 1 / wos;  cvisq^2 / wos
    } mc;
    int32_t plastic, dyn_bc, ell_w, n_probe;
    // dyn_bc jobs: 1 for each element block holding a node that some step
    // constrains (the blocks whose staged velocities the pipelined pass
    // overrides on step ticks); null otherwise
    const uint8_t* blk_bc;
    int32_t block_asm, pad_asm;  // 1: block assembly (staged mixed job, conflict-free blocks)
    uint32_t ne, nn, ne_s, nn_s;
    uint32_t nblk_e, nblk_n;
    uint32_t probe[MAX_PROBE];
    double* hist;
    long long hist_cap;
    // output snapshot ring (null when output is off): per slot the nodes'
    // x, a, v, u as output_results sees them (after compute_a, before the
    // velocity update), the elements' stress, and the time
    double* snap_node;  // [slots][nn][12]
    double* snap_elem;  // [slots][ne][6]
    double* snap_time;  // [slots]
    Ctl* ctl;
    double* part_e;  // [5][nblk_e]: min L, min dtStable, internal, hourglass, viscous work
    double* part_n;  // [2][nblk_n]: kinetic energy, external work of the step
    // element finalize spread over fin_g blocks (jobs with many block
    // partials): each block folds a fixed slice into fin_part[slice][0..6]
    // and counts itself on fin_ticket; the last one folds the slices in
    // slice order and resets the ticket
    double* fin_part;     // [fin_g][8]
    uint32_t* fin_ticket;
    uint32_t fin_g, pad_fin;
};

}  // namespace ensc
