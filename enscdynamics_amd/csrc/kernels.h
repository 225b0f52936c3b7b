// HIP kernels of the step engine (gfx950, wave64).
//
// One step iteration of the reference loop (exDyna3D.cpp:86-100) is one
// "tick" = three phases over every live job:
//   element pass : one element per lane.  Stages the element block's nodes
//               (X or x, v, u) in LDS (or gathers them from HBM), runs the
//               element body (elem_body.inc in FP64, elem_body_mixed.inc in
//               mixed precision), and writes either the element's 8 nodal
//               forces or, with block assembly, one partial force per
//               (element block, node); reduces min edge / min stable dt /
//               work terms per block.  k_element takes one element block
//               per launched block; k_element_pipe is a persistent,
//               software-pipelined version for block-assembly jobs.
//   k_fin_element : one block per job.  Closes the previous nodal update
//               (kinetic energy, external work, history row), reduces the
//               element pass's block partials in a fixed order and advances
//               the job's time-step scalars exactly as
//               update_minVertex_perMaterial_and_dt and the run-loop clamps
//               do (exDyna3D.cpp:663-686, 90-97, 775-780); decides the
//               output snapshot of the tick (output_results, 207-219).
//   k_node    : one node per lane.  Sums the node's forces (per-corner
//               forces in the canonical (corner, element) order, or the
//               block partials in block order), adds gravity, applies the
//               BCs and reactions (2029-2145), captures the output snapshot
//               when one is due, and does the central-difference update
//               a -> v -> u -> x (642-696).
//   k_fin_node : one block per job: closes the last nodal update of a run
//               of ticks (otherwise the next k_fin_element does).
// Off the step path: k_apply_bc (a step switch's velocity BCs) and k_post
// (the fields of an output snapshot).
// Cross-block results travel through per-block partials and the launch
// boundaries; no kernel waits on, or fences for, another block.
// The step size of a tick depends on the min edge of the current
// coordinates, which the element pass itself computes.  Only the overshoot
// clamp near the end time (682-685) needs that value before the pass; a tick
// that can hit it (decided from the previous pass's bound) runs a min-edge
// pre-pass instead and the next tick does the step.
#pragma once

#include <type_traits>

#include "engine_dev.h"
#include "../../include/ensc_engine.h"

namespace ensc {

// Rows of the reference's hourglass base vectors (exDyna3D.cpp:1003-1008).
__device__ constexpr double TAU_[4][8] = {{1., -1., 1., -1., 1., -1., 1., -1.},
                                          {-1., -1., 1., 1., 1., 1., -1., -1.},
                                          {-1., 1., 1., -1., 1., -1., -1., 1.},
                                          {-1., 1., -1., 1., 1., -1., 1., -1.}};

template <class R>
struct Vec4;
template <>
struct Vec4<double> {
    typedef double4 T;
};
template <>
struct Vec4<float> {
    typedef float4 T;
};

// Corner signs of the hex8 shape-function derivatives at the centre
// (Element_HexN8: dN(d, k) = CS[k][d] / 8).
__device__ constexpr double CS_[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                         {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

// Row k of the reference gradient, B0(k, i) = sum_j Ji(i,j) dN(j,k), from
// Gi = Ji / 8: each product is +-Gi(i,j) exactly, summed in the same order
// (j = 0, 1, 2) as the setup pass (fevalues.h:233-265), so the FP64 path
// reproduces the stored B0 bit for bit.
template <class R>
__device__ __forceinline__ void bgrad(const R Gi[3][3], int k, R Bk[3]) {
#pragma clang fp contract(off)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        R t = CS_[k][0] > 0 ? Gi[i][0] : -Gi[i][0];
        t += CS_[k][1] > 0 ? Gi[i][1] : -Gi[i][1];
        t += CS_[k][2] > 0 ? Gi[i][2] : -Gi[i][2];
        Bk[i] = t;
    }
}

// Walsh sign patterns of the 8 hex corners.  Pattern index s has bit 0 for
// xi, bit 1 for eta, bit 2 for zeta; its value at corner k is the product of
// the corner's signs CS_[k] over the axes in s.
enum : int { W_0 = 0, W_X = 1, W_Y = 2, W_XY = 3, W_Z = 4, W_XZ = 5, W_YZ = 6, W_XYZ = 7 };
// corner k -> binary index (xi bit + 2 eta bit + 4 zeta bit, bit set = +1)
__device__ constexpr int CORNER_BIN_[8] = {0, 1, 3, 2, 4, 5, 7, 6};

// w[s] = sum_k pattern_s(k) x[k]: a 3-stage butterfly (24 additions).
template <class R>
__device__ __forceinline__ void walsh8(const R x[8], R w[8]) {
    R a[8], b[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) a[CORNER_BIN_[k]] = x[k];
#pragma unroll
    for (int h = 1; h < 8; h <<= 1) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if (i & h) continue;
            b[i] = a[i | h] + a[i];      // axis not in the pattern
            b[i | h] = a[i | h] - a[i];  // axis in the pattern: (+1 side) - (-1 side)
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) a[i] = b[i];
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) w[i] = a[i];
}

// walsh8 of FP64 corner values evaluated in FP64 and rounded once to R at the
// end (the non-constant sums are differences of nearby values, so this keeps
// them exact to rounding whatever the common offset).
template <class R>
__device__ __forceinline__ void walsh8_d(const double c[8], R w[8]) {
    double wd[8];
    walsh8<double>(c, wd);
#pragma unroll
    for (int i = 0; i < 8; ++i) w[i] = R(wd[i]);
}

// f[k] = sum_s c[s] pattern_s(k): the expansion back to the corners.
template <class R>
__device__ __forceinline__ void walsh8_expand(const R c[8], R f[8]) {
    R a[8], b[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = c[i];
#pragma unroll
    for (int h = 1; h < 8; h <<= 1) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if (i & h) continue;
            b[i] = a[i] - a[i | h];      // -1 side of the axis
            b[i | h] = a[i] + a[i | h];  // +1 side
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) a[i] = b[i];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) f[k] = a[CORNER_BIN_[k]];
}

// An element's stored state, loaded at the top of the element pass (before
// the block's node staging, so the two sets of loads overlap).
template <class R>
struct ElemState {
    // gathered path: the corner node ids; staged path: the corners' LDS slots
    // in pairs, corners 2p (low 16 bits) and 2p + 1 (high) in nd[p], so the
    // pipelined pass prefetches them as they are stored and unpacks them at
    // use (NODE(k))
    uint32_t nd[8];
    R Gi[3][3];      // FP64 engine only (the mixed body rebuilds Gi and V0)
    R V0;
    R s[6];
    R Q[12];
    R ep;
};

template <class R, bool STAGED>
__device__ __forceinline__ void load_state(const JobDev& J, uint32_t e, bool full, ElemState<R>& S) {
    const uint32_t es = J.ne_s;
    if (STAGED) {
        const uint32_t* lc = reinterpret_cast<const uint32_t*>(J.lconn);
#pragma unroll
        for (int p = 0; p < 4; ++p) S.nd[p] = gld(lc, p * es + e);
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) S.nd[k] = gld(J.conn, k * es + e);
    }
    if (!full) return;
    const R* Sp = static_cast<const R*>(J.sigma);
    const R* Qp = static_cast<const R*>(J.qia);
    if (!std::is_same<R, float>::value) {
        // the mixed body rebuilds Gi and V0 from the reference positions
        const R* Gp = static_cast<const R*>(J.Gi);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) S.Gi[i][j] = gld(Gp, (3 * i + j) * es + e);
        S.V0 = gld(static_cast<const R*>(J.V0), e);
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) S.s[q] = gld(Sp, q * es + e);
#pragma unroll
    for (int q = 0; q < 12; ++q) S.Q[q] = gld(Qp, q * es + e);
    // elastic jobs read a valid word they never use (the bodies read ep only
    // for plastic jobs): an unconditional load, so the register is not first
    // zeroed, which would make the compiler wait for its previous load
    S.ep = gld(static_cast<const R*>(J.plastic ? J.epsp : J.sigma), e);
}

// Eigen 3.4's 3x3 determinant (Determinant.h:19-24, 56-65).
template <class R>
__device__ __forceinline__ R det3(const R m[3][3]) {
#pragma clang fp contract(off)
    const R h0 = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]);
    const R h1 = m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]);
    const R h2 = m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    return h0 - h1 + h2;
}

template <class R>
__device__ __forceinline__ R cof3(const R m[3][3], int i, int j) {
#pragma clang fp contract(off)
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return m[i1][j1] * m[i2][j2] - m[i1][j2] * m[i2][j1];
}

// Eigen 3.4's 3x3 cofactor inverse (InverseImpl.h:127-176).
template <class R>
__device__ __forceinline__ void inverse3(const R m[3][3], R r[3][3]) {
#pragma clang fp contract(off)
    const R c00 = cof3<R>(m, 0, 0), c10 = cof3<R>(m, 1, 0), c20 = cof3<R>(m, 2, 0);
    const R det = c00 * m[0][0] + c10 * m[1][0] + c20 * m[2][0];
    const R invdet = R(1) / det;
    r[1][0] = cof3<R>(m, 0, 1) * invdet;
    r[1][1] = cof3<R>(m, 1, 1) * invdet;
    r[2][0] = cof3<R>(m, 0, 2) * invdet;
    r[1][2] = cof3<R>(m, 2, 1) * invdet;
    r[2][1] = cof3<R>(m, 1, 2) * invdet;
    r[2][2] = cof3<R>(m, 2, 2) * invdet;
    r[0][0] = c00 * invdet;
    r[0][1] = c10 * invdet;
    r[0][2] = c20 * invdet;
}

// Reciprocal for the mixed-precision body: the hardware estimate (1 ulp)
// refined by one Newton step.  The bare estimate is not correctly rounded,
// and its one-sided errors add up over thousands of steps where the refined
// one does not.
__device__ __forceinline__ float rcp_nr(float x) {
    const float r = __builtin_amdgcn_rcpf(x);
    return __builtin_fmaf(r, __builtin_fmaf(-x, r, 1.0f), r);
}

// FP64 square root for the mixed-precision body's dt terms, whose inputs
// carry FP32 accuracy: the hardware reciprocal square root refined by one
// coupled Newton step for sqrt(x) and 1 / (2 sqrt(x)), then one correction of
// sqrt(x) (about 1 ulp; the library's correctly rounded root adds a second
// correction and range scaling).  Zero, infinity and NaN are returned as
// they are (the roots of 0 and +inf), like every other value outside the
// positive normals, which these inputs never are.
__device__ __forceinline__ double sqrt_nr(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double s = x * y, h = 0.5 * y;
    const double r = __builtin_fma(-s, h, 0.5);
    s = __builtin_fma(s, r, s);
    h = __builtin_fma(h, r, h);
    s = __builtin_fma(__builtin_fma(-s, s, x), h, s);
    return __builtin_amdgcn_class(x, 0x100) ? s : x;  // class bit 8: positive normal
}

// The cofactor inverse for the mixed-precision body: contracted products and
// one reciprocal (rcp_nr) in place of the IEEE division.  Returns the
// determinant (first-column expansion); *invdet gets its reciprocal.
__device__ __forceinline__ float inverse3_rcp(const float m[3][3], float r[3][3], float* invdet) {
#pragma clang fp contract(fast)
    auto cof = [&](int i, int j) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        return m[i1][j1] * m[i2][j2] - m[i1][j2] * m[i2][j1];
    };
    const float c00 = cof(0, 0), c10 = cof(1, 0), c20 = cof(2, 0);
    const float det = c00 * m[0][0] + c10 * m[1][0] + c20 * m[2][0];
    const float id = rcp_nr(det);
    r[0][0] = c00 * id;
    r[0][1] = c10 * id;
    r[0][2] = c20 * id;
    r[1][0] = cof(0, 1) * id;
    r[1][1] = cof(1, 1) * id;
    r[1][2] = cof(2, 1) * id;
    r[2][0] = cof(0, 2) * id;
    r[2][1] = cof(1, 2) * id;
    r[2][2] = cof(2, 2) * id;
    *invdet = id;
    return det;
}

// Compensation of the mixed engine's u sum (Kahan): per node one word of
// three signed 10-bit fields, field d the residual of the last addition to
// u[d] in units of 2^-10 ulp of u[d] (|residual| <= 1/2 ulp, so the fields
// keep it to 2^-11 ulp).
__device__ __forceinline__ double ucomp_get(uint32_t w, int d, double u) {
    const int q = ((int)(w << (22 - 10 * d))) >> 22;
    int k;
    (void)frexp(u, &k);  // ulp(u) = 2^(k - 53)
    return ldexp((double)q, k - 63);
}
__device__ __forceinline__ uint32_t ucomp_put(double c, double t, int d) {
    int k;
    (void)frexp(t, &k);
    const double s = fmin(fmax(rint(ldexp(c, 63 - k)), -512.0), 511.0);
    return ((uint32_t)(int)s & 0x3FFu) << (10 * d);
}

// Velocity-BC amplitude (extension): piecewise linear in total time, held
// constant outside the table.
__device__ inline double amp_value(const JobDev& J, int id, double t) {
#pragma clang fp contract(off)
    if (id < 0 || id >= J.n_amp) return 1.0;
    const int b = gld(J.amp_off, id), n = gld(J.amp_off, id + 1) - b;
    const ENSC_GLB double* T = (const ENSC_GLB double*)(J.amp_t + b);
    const ENSC_GLB double* A = (const ENSC_GLB double*)(J.amp_a + b);
    if (n <= 0) return 1.0;
    if (t <= T[0]) return A[0];
    if (t >= T[n - 1]) return A[n - 1];
    int i = 1;
    while (i < n && T[i] < t) ++i;
    const double w = (t - T[i - 1]) / (T[i] - T[i - 1]);
    return A[i - 1] + w * (A[i] - A[i - 1]);
}

// exDyna3D::deal_amp (exDyna3D.cpp:1222-1278): trapezoid mean of the
// amplitude over [time, time + dt].  Tables of fewer than 2 points give 1.
__device__ inline double deal_amp(const JobDev& J, int id, double time, double dt) {
#pragma clang fp contract(off)
    if (id < 0 || id >= J.n_amp) return 1.0;
    const int b = gld(J.amp_off, id), n = gld(J.amp_off, id + 1) - b;
    const ENSC_GLB double* T = (const ENSC_GLB double*)(J.amp_t + b);
    const ENSC_GLB double* A = (const ENSC_GLB double*)(J.amp_a + b);
    if (n < 2) return 1.0;
    double y = 1.0;
    for (int i = 0; i + 1 < n; ++i) {
        const double x1 = T[i], y1 = A[i], x2 = T[i + 1], y2 = A[i + 1];
        if (time > x2) break;
        const double m = (y2 - y1) / (x2 - x1);
        const double c = y1 - m * x1;
        const double xm = time + dt / 2.0;
        if ((xm + dt / 2.0) < x2) {
            y = m * xm + c;
            break;
        }
        const double DT1 = x2 - time;
        const double X = time + DT1 / 2.0;
        const double Y = m * X + c;
        const double A1 = Y * DT1;
        double A2 = 0.0;
        for (int j = 2; j + i < n; ++j) {
            const double xj = T[i + j], yj = A[i + j], xj_1 = T[i + j - 1], yj_1 = A[i + j - 1];
            if ((xm + dt / 2.0) > xj) {
                A2 += (xj - xj_1) * (yj + yj_1) / 2.0;
            } else {
                const double M = (yj - yj_1) / (xj - xj_1);
                const double C = yj_1 - m * xj_1;
                const double Y1 = M * (xm + dt / 2.0) + C;
                A2 += (xm + dt / 2.0 - xj_1) * (yj_1 + Y1) / 2.0;
                break;
            }
        }
        const double x_end = T[n - 1];
        if (xm + dt / 2.0 > x_end) A2 += (xm + dt / 2.0 - x_end) * 1.0;
        y = (A1 + A2) / dt;
        break;
    }
    return y;
}

// A job's scalars as one uniform read through the constant address space,
// so the compiler issues scalar loads together and waits for them alone.
// Read through generic pointers the fields came one at a time, each load
// followed by a wait for every outstanding vector memory operation of the
// wave (in the pipelined element pass: the tile's copies and the previous
// tile's stores).  Only for kernels that do not write the scalars they read
// (the element kernels set the sticky error word by atomics; nothing here
// reads it): the scalars come from an earlier launch, and a launch starts
// with the scalar cache invalidated.  Used where it measured faster: the
// pipelined element pass (once per tile) and the nodal update.
__device__ __forceinline__ Ctl ctl_read(const Ctl* p) {
    static_assert(sizeof(Ctl) % 4 == 0, "Ctl is read as 32-bit words");
    typedef const __attribute__((address_space(4))) uint32_t* KWord;
    const KWord q = (KWord)p;
    Ctl c;
    uint32_t* d = reinterpret_cast<uint32_t*>(&c);
#pragma unroll
    for (unsigned i = 0; i < sizeof(Ctl) / 4; ++i) d[i] = q[i];
    return c;
}

struct Decision {
    int32_t kind;
    double dt_i_1;  // the increment the element pass integrates with
    double time_n;  // time after this step's increment
};

// What this tick does for one job, from its scalars at tick start.  Every
// block of the job evaluates it identically.
__device__ inline Decision decide(const Ctl& c, double factor, double wos) {
#pragma clang fp contract(off)
    Decision d{KIND_NONE, 0.0, c.time};
    if (c.init_pending) {
        d.kind = KIND_INIT;
        return d;
    }
    if (!(c.time < c.step_end) || c.nsteps >= c.step_limit) return d;
    // every output slot holds an undrained snapshot: wait for the host
    if (c.out_on && c.snap_head - c.snap_tail >= c.snap_cap) return d;
    double dt1 = c.dt_i;
    const double rem = c.total_time - (c.time + dt1);
    // dt_i <= dtStable_min * factor always, so the overshoot clamp of
    // exDyna3D.cpp:682-685 can only fire inside this band.
    const bool band = (rem > 0.0) && (rem < c.dts_min * factor);
    if (band && !c.have_lpre) {
        d.kind = KIND_PREPASS;
        return d;
    }
    if (band) {
        double dt = c.lmin_pre / wos;
        const double tmp = c.dts_min;
        dt = dt < tmp ? dt : tmp;
        dt *= factor;
        if (rem < dt && rem > 0.0) dt1 = rem;
    }
    if (c.time + dt1 > c.step_end) dt1 = c.step_end - c.time;
    d.kind = KIND_STEP;
    d.dt_i_1 = dt1;
    double t = c.time + dt1;
    if (t > c.total_time) t = c.total_time;
    d.time_n = t;
    return d;
}

// Block barrier for LDS traffic only.  __syncthreads() (and a workgroup
// fence on LDS, which also covers global->LDS copies) makes every wave wait
// for all its outstanding global loads, stores and copies; the element
// kernels exchange data between waves through LDS written by ordinary LDS
// stores, so they only need those to complete: s_waitcnt lgkmcnt(0)
// (immediate 0xC07F: vmcnt and expcnt left at their maximum), then the
// barrier.  The signal fences keep the compiler from moving memory
// accesses across it.
__device__ __forceinline__ void lds_sync() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Wave reductions of doubles through DPP row moves (no LDS): row shifts by
// 1, 2, 4 and 8 leave each 16-lane row's total in its lane 15, then
// row_bcast:15 and row_bcast:31 carry it across rows, so lane 63 holds the
// wave's total (min or sum).  Lane 63's value depends only on lanes whose
// source lies inside their row at every step (15 <- 14, 13, 11, 7 and
// their own chains), so the moves need no identity: a lane without a
// source reads 0 (bound control) or keeps its register, and its result,
// wrong, feeds no lane that lane 63 reads.  The order of the operations is
// fixed, so the result is too.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move(double v) {
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)x, CTRL, ROWS, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(x >> 32), CTRL, ROWS, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <bool MIN>
__device__ __forceinline__ double dpp_op(double a, double b) {
    return MIN ? __builtin_fmin(a, b) : a + b;
}
template <bool MIN>
__device__ __forceinline__ double wave_reduce63(double v) {
    v = dpp_op<MIN>(v, dpp_move<0x111, 0xF>(v));  // row_shr:1
    v = dpp_op<MIN>(v, dpp_move<0x112, 0xF>(v));  // row_shr:2
    v = dpp_op<MIN>(v, dpp_move<0x114, 0xF>(v));  // row_shr:4
    v = dpp_op<MIN>(v, dpp_move<0x118, 0xF>(v));  // row_shr:8
    v = dpp_op<MIN>(v, dpp_move<0x142, 0xA>(v));  // row_bcast:15 into rows 1, 3
    v = dpp_op<MIN>(v, dpp_move<0x143, 0xC>(v));  // row_bcast:31 into rows 2, 3
    return v;                                      // the wave's value in lane 63
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <class T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const T w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}

// Velocity BCs with an amplitude (extension) change the prescribed velocity
// every step; the element pass applies the value at this step's time to the
// corners it reads.  Out of line: jobs without amplitudes never call it.
// Values travel by value so the caller keeps them in registers.
struct V3 {
    double x, y, z;
};
__device__ __noinline__ double bc_override_c(const JobDev& J, uint32_t g, double t, int d, double v) {
#pragma clang fp contract(off)
    const uint8_t mk = gld(J.bcmask, g);
    if (mk & (BC_VEL0 << d)) v = gld(J.bcval, g * 3 + d) * amp_value(J, gld(J.bcamp, g * 3 + d), t);
    if (mk & (BC_SPC0 << d)) v = 0.0;
    return v;
}
__device__ __forceinline__ V3 bc_override_inl(const JobDev& J, uint32_t g, double t, V3 in) {
#pragma clang fp contract(off)
    const uint8_t mk = gld(J.bcmask, g);
    if (!mk) return in;
    double v[3] = {in.x, in.y, in.z};
    for (int d = 0; d < 3; ++d) {
        if (mk & (BC_VEL0 << d)) v[d] = gld(J.bcval, g * 3 + d) * amp_value(J, gld(J.bcamp, g * 3 + d), t);
        if (mk & (BC_SPC0 << d)) v[d] = 0.0;
    }
    return V3{v[0], v[1], v[2]};
}
__device__ __noinline__ V3 bc_override(const JobDev& J, uint32_t g, double t, V3 in) {
    return bc_override_inl(J, g, t, in);
}

#define CORNER_FENCE
// Element body instantiations: corner data gathered from HBM (G) or from
// the block's node records staged in LDS (S).  NVAL(k, i) is value i of
// corner k's x0 x1 x2 v0 v1 v2 u0 u1 u2, with x the current position (in
// the mixed engine the record holds the reference position, so x = X + u).

#define ACC_W 3                  // block assembly: 3 floats per staged node
#define ASM_BEGIN ((void)0)      // the accumulators start zeroed
#define PREFETCH_HOOK ((void)0)
#define ELEM_ARGS                                                                                          \
    const JobDev &J, uint32_t e, const ElemState<R> &ES, double dt, int32_t kind, double time_n, bool act, \
        bool asmode, double &Le, double &dts, double *en, uint32_t &errf, const ENSC_LDS double *lds,     \
        ENSC_LDS R *acc
#define NODE(k) ES.nd[k]
#define FIRST(k) 0u  // gathered corners: no block assembly
#define NVAL(k, i)                                                                                      \
    ((i) < 3 ? (MIXED ? gld((const double*)J.pos, NODE(k) * POS_REC + (i)) +                           \
                            gld((const double*)J.vu, NODE(k) * VU_REC + 3 + (i))                           \
                      : gld((const double*)J.pos, NODE(k) * POS_REC + (i)))                              \
             : gld((const double*)J.vu, NODE(k) * VU_REC + (i)-3))
#define DYN_BC_OVERRIDE(k, v)                                              \
    if (J.dyn_bc && kind == KIND_STEP) {                                   \
        const V3 o_ = bc_override(J, NODE(k), time_n, V3{v[0], v[1], v[2]}); \
        v[0] = o_.x;                                                       \
        v[1] = o_.y;                                                       \
        v[2] = o_.z;                                                       \
    }
#define CORNER_V(k, i) \
    ((J.dyn_bc && kind == KIND_STEP) ? bc_override_c(J, NODE(k), time_n, (i), NVAL(k, 3 + (i))) : NVAL(k, 3 + (i)))
template <class R>
__device__ __forceinline__ void element_fp64_G(ELEM_ARGS) {
#pragma clang fp contract(off)
    constexpr bool MIXED = false, ASM_CAP = false;
#include "elem_body.inc"
}
template <class R>
__device__ __forceinline__ void element_mixed_G(ELEM_ARGS) {
#pragma clang fp contract(fast)
    constexpr bool MIXED = true, ASM_CAP = false;
#include "elem_body_mixed.inc"
}
#undef NODE
#undef FIRST
#undef NVAL
#undef DYN_BC_OVERRIDE
#undef CORNER_V

// Staged record in LDS: [x0 x1 x2 v0 v1 v2 u0 u1 u2] as copied from the node
// record (x0..2 = the record's X).  A corner's 16-bit slot carries in bit 15
// whether it is the first contribution to its node in the assembly's phase
// order (FIRST(k)).
#define NODE(k) (((k)&1) ? ((ES.nd[(k) >> 1] >> 16) & 0x7FFFu) : (ES.nd[(k) >> 1] & 0x7FFFu))
#define FIRST(k) (((k)&1) ? (ES.nd[(k) >> 1] >> 31) : ((ES.nd[(k) >> 1] >> 15) & 1u))
#define NVAL(k, i) \
    ((MIXED && (i) < 3) ? lds[NODE(k) * 9 + (i)] + lds[NODE(k) * 9 + 6 + (i)] : lds[NODE(k) * 9 + (i)])
#define DYN_BC_OVERRIDE(k, v)  // applied once per node while staging
#define CORNER_V(k, i) NVAL(k, 3 + (i))
template <class R>
__device__ __forceinline__ void element_fp64_S(ELEM_ARGS) {
#pragma clang fp contract(off)
    constexpr bool MIXED = false, ASM_CAP = true;
#include "elem_body.inc"
}
template <class R>
__device__ __forceinline__ void element_mixed_S(ELEM_ARGS) {
#pragma clang fp contract(fast)
    constexpr bool MIXED = true, ASM_CAP = true;
#include "elem_body_mixed.inc"
}
// FP64 block-assembly jobs: the Walsh-form FP64 body
template <class R>
__device__ __forceinline__ void element_fp64_W(ELEM_ARGS) {
#pragma clang fp contract(fast)
    constexpr bool MIXED = false, ASM_CAP = true;
#include "elem_body_walsh64.inc"
}
#undef NODE
#undef FIRST
#undef NVAL
#undef DYN_BC_OVERRIDE
#undef CORNER_V
#undef ELEM_ARGS
#undef ACC_W
#undef ASM_BEGIN
#undef PREFETCH_HOOK

// Fixed-order reduction of a job's block partials (value q of block b at
// part[q * nblk + b]; values 0-1 are the node partials, 2-6 the element
// partials).  fold_partials: thread t folds blocks t, t+BLK, t+2*BLK, ... in
// ascending order, for the values in VM (bit q: value q; the others keep
// their identity), loads issued FOLD_U blocks at a time (one round for jobs
// of up to FOLD_U * BLK blocks).  It reads only the job table, so its loads
// go out together with those of the job's scalars; sums of a phase that is
// not open are not used.  combine_partials: a fixed DPP reduction per wave
// (wave_reduce63) of the values in VM, then thread 0 adds the waves in index
// order.  Same inputs, same result.  Every thread of the block must call
// both.
constexpr int FOLD_U = 8;  // blocks per thread whose loads are issued together
__host__ __device__ constexpr bool fold_min(int q) { return q == 2 || q == 3; }  // values 2-3 are minima
// The slice [slice, slice + 1) of `slices` equal parts of the partials (in
// ascending block order) is folded; one slice folds them all.
template <int BLK, uint32_t VM>
__device__ inline void fold_partials(const JobDev& J, double* acc, uint32_t slice = 0, uint32_t slices = 1) {
    const uint32_t t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < 7; ++q) acc[q] = fold_min(q) ? __builtin_inf() : 0.0;
    const uint32_t nn_all = (VM & 3u) ? J.nblk_n : 0u, ne_all = (VM & 0x7Cu) ? J.nblk_e : 0u;
    const uint32_t sn = (nn_all + slices - 1) / slices, se = (ne_all + slices - 1) / slices;
    const uint32_t n0 = sn * slice, e0 = se * slice;
    const uint32_t nn = nn_all > n0 ? (nn_all - n0 < sn ? nn_all - n0 : sn) : 0u;  // this slice's counts
    const uint32_t ne = ne_all > e0 ? (ne_all - e0 < se ? ne_all - e0 : se) : 0u;
    const ENSC_GLB double* pn = (const ENSC_GLB double*)J.part_n;
    const ENSC_GLB double* pe = (const ENSC_GLB double*)J.part_e;
    const uint32_t nmax = nn > ne ? nn : ne;
    for (uint32_t b0 = t; b0 < nmax; b0 += FOLD_U * BLK) {
        double v[FOLD_U][7];
        // every load reads a valid partial (index clamped), so they issue
        // together without branches; out-of-range ones are replaced after
#pragma unroll
        for (int u = 0; u < FOLD_U; ++u) {
            const uint32_t b = b0 + u * BLK;
            const uint32_t bn = n0 + (nn ? (b < nn ? b : nn - 1) : 0u), be = e0 + (ne ? (b < ne ? b : ne - 1) : 0u);
            const uint32_t bn_ok = nn_all ? bn : 0u, be_ok = ne_all ? be : 0u;
#pragma unroll
            for (int q = 0; q < 7; ++q) {
                if (!((VM >> q) & 1u)) {
                    v[u][q] = fold_min(q) ? __builtin_inf() : 0.0;
                } else if (q < 2) {
                    v[u][q] = pn[(size_t)q * nn_all + (nn_all ? (bn_ok < nn_all ? bn_ok : nn_all - 1) : 0u)];
                    v[u][q] = (b < nn) ? v[u][q] : 0.0;
                } else {
                    v[u][q] = pe[(size_t)(q - 2) * ne_all + (ne_all ? (be_ok < ne_all ? be_ok : ne_all - 1) : 0u)];
                    v[u][q] = (b < ne) ? v[u][q] : (fold_min(q) ? __builtin_inf() : 0.0);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < FOLD_U; ++u)
#pragma unroll
            for (int q = 0; q < 7; ++q)
                if ((VM >> q) & 1u) acc[q] = fold_min(q) ? (v[u][q] < acc[q] ? v[u][q] : acc[q]) : acc[q] + v[u][q];
    }
}

template <int BLK, uint32_t VM>
__device__ inline void combine_partials(double* acc, double (*s_tmp)[7], double* out) {
    const uint32_t t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < 7; ++q)
        if ((VM >> q) & 1u) acc[q] = fold_min(q) ? wave_reduce63<true>(acc[q]) : wave_reduce63<false>(acc[q]);
    const int w = t >> 6;
    if ((t & 63) == 63)
#pragma unroll
        for (int q = 0; q < 7; ++q)
            if ((VM >> q) & 1u) s_tmp[w][q] = acc[q];
    __syncthreads();
    if (t == 0) {
#pragma unroll
        for (int q = 0; q < 7; ++q) {
            if (!((VM >> q) & 1u)) {
                out[q] = fold_min(q) ? __builtin_inf() : 0.0;
                continue;
            }
            double r = s_tmp[0][q];
            for (int x = 1; x < BLK / 64; ++x) r = fold_min(q) ? (s_tmp[x][q] < r ? s_tmp[x][q] : r) : r + s_tmp[x][q];
            out[q] = r;
        }
    }
}

// Row `row` of the job's history, or null (no history, or past its end).
__device__ __forceinline__ double* hist_row(const JobDev& J, long long row) {
    return (J.hist && row >= 0 && row < J.hist_cap) ? J.hist + row * (ENSC_H_BASE + 3 * J.n_probe) : nullptr;
}

// End of a job's nodal update (thread 0): kinetic energy, external work (w:
// the step's work summed over the nodes) and the job's history row for the
// step; marks the update closed.  c is the
// job's Ctl or the caller's register copy of it (written back once): updating
// the device copy field by field makes every later read of it wait for the
// stores before it, which the per-tick finalize avoids.
// A history row's time-step columns and probe displacements.
__device__ __forceinline__ void hist_time_columns(const JobDev& J, const Ctl& c, double* h) {
    h[ENSC_H_TIME] = c.time;
    h[ENSC_H_DT_I] = c.dt_i;
    h[ENSC_H_DT_I_1] = c.dt_i_1;
    h[ENSC_H_DTS_MIN] = c.dts_min;
    for (int p = 0; p < J.n_probe; ++p) {
        const double* r = J.vu + (size_t)J.probe[p] * VU_REC + 3;
        h[ENSC_H_BASE + 3 * p + 0] = r[0];
        h[ENSC_H_BASE + 3 * p + 1] = r[1];
        h[ENSC_H_BASE + 3 * p + 2] = r[2];
    }
}
__device__ __forceinline__ void node_phase_finish(const JobDev& J, Ctl& c, double ke, double w) {
#pragma clang fp contract(off)
    c.ke = ke;
    c.wext += w;
    if (double* h = hist_row(J, c.nsteps - 1)) {
        hist_time_columns(J, c, h);
        h[ENSC_H_KE] = c.ke;
        h[ENSC_H_IE] = c.ie;
        h[ENSC_H_HG] = c.hg;
        h[ENSC_H_VIS] = c.vis;
        h[ENSC_H_WEXT] = c.wext;
    }
    c.kind = KIND_NONE;
    c.tick_nsteps = 0;
}

// End of a job's element pass (thread 0): tot = min edge, min dtStable,
// stress, hourglass and viscous work of the pass.  c as in node_phase_finish.
__device__ __forceinline__ void elem_phase_finish(const JobDev& J, Ctl& c, const Decision& dc, const double* tot) {
#pragma clang fp contract(off)
    const double Lmin = tot[0], dtsn = tot[1];
    c.snap_now = 0;
    if (dc.kind == KIND_PREPASS) {
        c.lmin_pre = Lmin;
        c.have_lpre = 1;
        c.kind = KIND_PREPASS;
    } else if (dc.kind == KIND_INIT) {
        c.dts_min = dtsn;
        c.init_pending = 0;
        c.kind = KIND_INIT;
    } else {
        // update_minVertex_perMaterial_and_dt (663-686) + run-loop clamp (90-97)
        double dt1 = c.dt_i;
        const double time_prev = c.time + dt1;
        double dt = Lmin / J.wos;
        const double tmp = c.dts_min;
        dt = dt < tmp ? dt : tmp;
        dt *= J.factor;
        if ((c.total_time - time_prev) < dt && (c.total_time - time_prev) > 0.0) {
            dt = c.total_time - time_prev;
            dt1 = dt;
        }
        if (c.time + dt1 > c.step_end) dt = dt1 = c.step_end - c.time;
        if (dt1 != dc.dt_i_1) c.err |= (uint32_t)ENSC_EF_DT;  // cannot happen: same inputs as decide()
        // computeSate (775-780)
        double t = c.time + dt1;
        if (t > c.total_time) t = c.total_time;
        c.time = t;
        c.dt_i = dt;
        c.dt_i_1 = dt1;
        c.dts_min = dtsn;
        c.nsteps += 1;
        c.have_lpre = 0;
        c.ie += tot[2];
        c.hg += tot[3];
        c.vis += tot[4];
        c.grav_amp = (J.grav_on && J.grav_amp >= 0) ? deal_amp(J, J.grav_amp, t, dt) : 1.0;
        if (!(dtsn > 0.0) || !(dt >= 0.0)) c.err |= (uint32_t)ENSC_EF_DT;
        // output_results (207-219): this step's state is written when time
        // reaches the next output time or the total time
        if (c.out_on && (t >= c.t_out || t >= c.total_time)) {
            const int32_t slot = c.snap_head % c.snap_cap;
            c.snap_now = slot + 1;
            c.snap_head += 1;
            c.t_out += c.t_int;
            J.snap_time[slot] = t;
        }
        c.kind = KIND_STEP;
    }
}

// W64: FP64 block-assembly jobs only (staged, FP64), with the Walsh-form body
// on every tick (a pre-pass tick computes the min edge only).
template <bool MIXED, int BLK, bool STAGED, int WPE, bool W64 = false>
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(WPE))) void k_element(const JobDev* __restrict__ jobs, const uint4* __restrict__ blkmap) {
#pragma clang fp contract(off)
    extern __shared__ double s_nodes[];  // STAGED: [block's nodes][x0 x1 x2 v0 v1 v2 u0 u1 u2]
    // (job, block of the job, first position of the block's node list, its length)
    const uint4 bm = blkmap[blockIdx.x];
    const JobDev& J = jobs[bm.x];
    Ctl* ctl = J.ctl;
    const uint32_t lb = bm.y, n0 = bm.z, nbn = bm.w;
    const uint32_t e = lb * BLK + threadIdx.x;
    const bool act = e < J.ne;
    typedef typename std::conditional<MIXED, float, double>::type R;
    // generic loads here: with one element block per launched block the
    // scalar read measured slower (FP64 element pass 128 -> 133 us)
    const Decision dc = decide(*ctl, J.factor, J.wos);
    if (dc.kind == KIND_NONE) {
        if (lb == 0 && threadIdx.x == 0) gst(&ctl->tick_elem, 0u) = KIND_NONE;  // the tick record
        return;
    }
    ElemState<R> ES;
    load_state<R, STAGED>(J, act ? e : J.ne - 1, dc.kind != KIND_PREPASS, ES);
    ENSC_LDS double* lds = (ENSC_LDS double*)s_nodes;
    // block assembly: staged jobs whose blocks are conflict-free, on ticks
    // that compute forces (uniform over the block)
    const bool asmode = STAGED && J.block_asm && dc.kind != KIND_PREPASS;
    ENSC_LDS R* acc = (ENSC_LDS R*)(lds + nbn * 9);  // block assembly: sum per block node, precision R
    if (STAGED) {
        // stage the block's node records once (X, v, u); velocity BCs with an
        // amplitude are applied here, once per node
        for (uint32_t i = threadIdx.x; i < nbn; i += BLK) {
            const uint32_t g = gld(J.blk_nodes, n0 + i);
            const D2 p0 = gld2(J.pos, g * POS_REC), p1 = gld2(J.pos, g * POS_REC + 2);
            const D2 p2 = gld2(J.vu, g * VU_REC), p3 = gld2(J.vu, g * VU_REC + 2), p4 = gld2(J.vu, g * VU_REC + 4);
            // [X0 X1 | X2 im] [v0 v1 | v2 u0 | u1 u2]
            V3 v{p2.a, p2.b, p3.a};
            if (J.dyn_bc && dc.kind == KIND_STEP) v = bc_override(J, g, dc.time_n, v);
            ENSC_LDS double* d = lds + i * 9;
            d[0] = p0.a;
            d[1] = p0.b;
            d[2] = p1.a;
            d[3] = v.x;
            d[4] = v.y;
            d[5] = v.z;
            d[6] = p3.b;
            d[7] = p4.a;
            d[8] = p4.b;
            if (asmode) {
                acc[3 * i] = R(0);
                acc[3 * i + 1] = R(0);
                acc[3 * i + 2] = R(0);
            }
        }
        lds_sync();
    }
    double Le = __builtin_inf(), dts = __builtin_inf();
    double en[3] = {0.0, 0.0, 0.0};
    uint32_t errf = 0;
    if constexpr (W64) {
        static_assert(STAGED && !MIXED, "the Walsh-form FP64 body stages its nodes");
        // every lane runs the body (see below); a pre-pass stops at the min edge
        if (asmode || act)
            element_fp64_W<R>(J, act ? e : J.ne - 1, ES, dc.dt_i_1, dc.kind, dc.time_n, act, asmode, Le, dts, en, errf,
                              lds, acc);
        if (asmode) {
            lds_sync();
            for (uint32_t i = threadIdx.x; i < nbn; i += BLK) {
                R* pf = static_cast<R*>(J.part_f) + (size_t)(n0 + i) * 3;
                gst(pf, 0u) = acc[3 * i];
                gst(pf, 1u) = acc[3 * i + 1];
                gst(pf, 2u) = acc[3 * i + 2];
            }
        }
        if (!act) {
            Le = dts = __builtin_inf();
            en[0] = en[1] = en[2] = 0.0;
        }
    } else if (asmode) {
        // every lane runs the body (padding lanes on the last element) so the
        // whole block meets the assembly barriers; padding results are dropped
        if constexpr (MIXED && STAGED)
            element_mixed_S<R>(J, act ? e : J.ne - 1, ES, dc.dt_i_1, dc.kind, dc.time_n, act, true, Le, dts, en, errf,
                               lds, acc);
        else if constexpr (STAGED)
            element_fp64_S<R>(J, act ? e : J.ne - 1, ES, dc.dt_i_1, dc.kind, dc.time_n, act, true, Le, dts, en, errf,
                              lds, acc);
        lds_sync();
        for (uint32_t i = threadIdx.x; i < nbn; i += BLK) {
            R* pf = static_cast<R*>(J.part_f) + (size_t)(n0 + i) * 3;
            gst(pf, 0u) = acc[3 * i];
            gst(pf, 1u) = acc[3 * i + 1];
            gst(pf, 2u) = acc[3 * i + 2];
        }
        if (!act) {
            Le = dts = __builtin_inf();
            en[0] = en[1] = en[2] = 0.0;
        }
    } else if (act) {
        if constexpr (STAGED) {
            if constexpr (MIXED)
                element_mixed_S<R>(J, e, ES, dc.dt_i_1, dc.kind, dc.time_n, true, false, Le, dts, en, errf, lds, acc);
            else
                element_fp64_S<R>(J, e, ES, dc.dt_i_1, dc.kind, dc.time_n, true, false, Le, dts, en, errf, lds, acc);
        } else {
            if constexpr (MIXED)
                element_mixed_G<R>(J, e, ES, dc.dt_i_1, dc.kind, dc.time_n, true, false, Le, dts, en, errf, lds, acc);
            else
                element_fp64_G<R>(J, e, ES, dc.dt_i_1, dc.kind, dc.time_n, true, false, Le, dts, en, errf, lds, acc);
        }
    }
    // block partials: wave shuffles, then LDS across the block's waves
    __shared__ double s_red[BLK / 64][5];
    __shared__ uint32_t s_err;
    if (threadIdx.x == 0) s_err = 0;
    Le = wave_reduce63<true>(Le);
    dts = wave_reduce63<true>(dts);
    en[0] = wave_reduce63<false>(en[0]);
    en[1] = wave_reduce63<false>(en[1]);
    en[2] = wave_reduce63<false>(en[2]);
    lds_sync();
    if (errf) atomicOr(&s_err, errf);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) {
        s_red[w][0] = Le;
        s_red[w][1] = dts;
        s_red[w][2] = en[0];
        s_red[w][3] = en[1];
        s_red[w][4] = en[2];
    }
    lds_sync();
    if (threadIdx.x == 0) {
        double r0 = s_red[0][0], r1 = s_red[0][1], r2 = s_red[0][2], r3 = s_red[0][3], r4 = s_red[0][4];
        for (int q = 1; q < BLK / 64; ++q) {
            r0 = s_red[q][0] < r0 ? s_red[q][0] : r0;
            r1 = s_red[q][1] < r1 ? s_red[q][1] : r1;
            r2 += s_red[q][2];
            r3 += s_red[q][3];
            r4 += s_red[q][4];
        }
        const uint32_t nb = J.nblk_e;
        gst(J.part_e, lb) = r0;
        gst(J.part_e, nb + lb) = r1;
        gst(J.part_e, 2 * nb + lb) = r2;
        gst(J.part_e, 3 * nb + lb) = r3;
        gst(J.part_e, 4 * nb + lb) = r4;
        if (s_err) atomicOr(&ctl->err, s_err);
        if (lb == 0) gst(&ctl->tick_elem, 0u) = dc.kind;  // the tick record
    }
}

// Per-job end of the element pass.  It closes the previous nodal update if
// still open (node_phase_finish), and reduces the job's element block
// partials (written by k_element; the launch boundary orders them) to
// advance the job's time-step scalars exactly as
// update_minVertex_perMaterial_and_dt and the run-loop clamps do
// (exDyna3D.cpp:663-686, 90-97, 775-780).
//
// The map holds (job, slice, slices, role) per block.  A job with more than
// 1024 and up to 4096 partials per set has four blocks, which run side by
// side, each folding the partials of the scalars only it writes (one CU's
// loads bound the fold, so four CUs fold a quarter each):
//   FIN_DECIDE: min edge and min stable dt -> the decision, the time-step
//     scalars and the history row's time columns;
//   FIN_NODE:   kinetic energy and the step's external work of the nodal
//     update being closed -> ke, wext and their history columns;
//   FIN_WORK_A / FIN_WORK_B: internal and hourglass work / viscous work ->
//     ie, hg / vis and their history columns (the values before this
//     pass's work, as node_phase_finish records them).
// The deciding block rewrites kind and nsteps, so the other three learn
// what the tick did from the tick record (tick_nsteps, tick_elem), written
// by the nodal update and the element pass before this launch; each block
// stores its own fields only.
// Other jobs have fin_g blocks of role FIN_ALL: one for a small job, which
// folds all seven values and finishes both phases; for a job of more than
// 4096 partials each folds its slice of all seven values, publishes it and
// counts itself on the job's ticket (device-scope fences around the
// atomic); the block that counts last folds the slices in slice order, so
// the result does not depend on which block that is, and finishes both
// phases.
enum : uint32_t { FIN_ALL = 0, FIN_DECIDE = 1, FIN_NODE = 2, FIN_WORK_A = 3, FIN_WORK_B = 4 };
template <int BLK>
__global__ __launch_bounds__(BLK) void k_fin_element(const JobDev* __restrict__ jobs,
                                                     const uint4* __restrict__ finmap) {
#pragma clang fp contract(off)
    const uint4 fm = finmap[blockIdx.x];  // (job, slice, slices, role)
    const JobDev& J = jobs[fm.x];
    const uint32_t slice = fm.y, slices = fm.z, role = fm.w;
    Ctl* ctl = J.ctl;
    // one copy of the job's scalars, loaded at once.  FIN_ALL: nothing else
    // writes them while this kernel runs (the job's deciding block stores
    // them once every block has counted itself, i.e. after every block has
    // read them).  The roles: each reads only fields that no other block of
    // the job writes in this launch.
    const Ctl c0 = *ctl;
    __shared__ double s_red[BLK / 64][7];
    double acc[7], tot[7];
    if (role == FIN_NODE) {
        if (!c0.tick_nsteps) return;
        fold_partials<BLK, 0x03u>(J, acc);
        combine_partials<BLK, 0x03u>(acc, s_red, tot);
        if (threadIdx.x == 0) {
            const double ke = tot[0], wext = c0.wext + tot[1];
            ctl->ke = ke;
            ctl->wext = wext;
            if (double* h = hist_row(J, c0.tick_nsteps - 1)) {
                h[ENSC_H_KE] = ke;
                h[ENSC_H_WEXT] = wext;
            }
        }
        return;
    }
    if (role == FIN_WORK_A || role == FIN_WORK_B) {
        const bool step = c0.tick_elem == KIND_STEP;
        if (!step && !c0.tick_nsteps) return;
        if (role == FIN_WORK_A) {
            fold_partials<BLK, 0x30u>(J, acc);
            combine_partials<BLK, 0x30u>(acc, s_red, tot);
        } else {
            fold_partials<BLK, 0x40u>(J, acc);
            combine_partials<BLK, 0x40u>(acc, s_red, tot);
        }
        if (threadIdx.x == 0) {
            double* h = c0.tick_nsteps ? hist_row(J, c0.tick_nsteps - 1) : nullptr;
            if (role == FIN_WORK_A) {
                if (h) {
                    h[ENSC_H_IE] = c0.ie;
                    h[ENSC_H_HG] = c0.hg;
                }
                if (step) {
                    ctl->ie = c0.ie + tot[4];
                    ctl->hg = c0.hg + tot[5];
                }
            } else {
                if (h) h[ENSC_H_VIS] = c0.vis;
                if (step) ctl->vis = c0.vis + tot[6];
            }
        }
        return;
    }
    const bool node_open = c0.kind == KIND_STEP;
    const Decision dc = decide(c0, J.factor, J.wos);
    const bool elem = dc.kind != KIND_NONE;
    if (role == FIN_DECIDE) {
        if (!node_open && !elem) return;
        fold_partials<BLK, 0x0Cu>(J, acc);
        combine_partials<BLK, 0x0Cu>(acc, s_red, tot);
        if (threadIdx.x == 0) {
            Ctl c = c0;
            if (node_open) {
                // node_phase_finish's columns of this block's fields
                if (double* h = hist_row(J, c.nsteps - 1)) hist_time_columns(J, c, h);
                c.kind = KIND_NONE;
            }
            // the work sums (tot[4..6]) are the identity here and ie, hg,
            // vis are not stored: FIN_WORK_A and FIN_WORK_B own them
            if (elem) elem_phase_finish(J, c, dc, tot + 2);
            ctl->time = c.time;
            ctl->dt_i = c.dt_i;
            ctl->dt_i_1 = c.dt_i_1;
            ctl->dts_min = c.dts_min;
            ctl->lmin_pre = c.lmin_pre;
            ctl->grav_amp = c.grav_amp;
            ctl->nsteps = c.nsteps;
            ctl->kind = c.kind;
            ctl->have_lpre = c.have_lpre;
            ctl->init_pending = c.init_pending;
            ctl->err = c.err;
            ctl->t_out = c.t_out;
            ctl->snap_now = c.snap_now;
            ctl->snap_head = c.snap_head;
        }
        return;
    }
    // FIN_ALL
    fold_partials<BLK, 0x7Fu>(J, acc, slice, slices);
    if (!node_open && !elem) return;  // alike in every block of the job
    combine_partials<BLK, 0x7Fu>(acc, s_red, tot);
    if (slices > 1) {
        __shared__ uint32_t s_last;
        if (threadIdx.x == 0) {
            ENSC_GLB double* fp = (ENSC_GLB double*)J.fin_part;
#pragma unroll
            for (int q = 0; q < 7; ++q) fp[slice * 8 + q] = tot[q];
            __threadfence();  // the slice before the count
            s_last = atomicAdd(J.fin_ticket, 1u) == slices - 1 ? 1u : 0u;
        }
        __syncthreads();
        if (!s_last) return;
        if (threadIdx.x == 0) {
            __threadfence();  // every counted slice after the count
            const ENSC_GLB double* fp = (const ENSC_GLB double*)J.fin_part;
#pragma unroll
            for (int q = 0; q < 7; ++q) tot[q] = fp[q];
            for (uint32_t g = 1; g < slices; ++g)
#pragma unroll
                for (int q = 0; q < 7; ++q) {
                    const double x = fp[g * 8 + q];
                    tot[q] = fold_min(q) ? (x < tot[q] ? x : tot[q]) : tot[q] + x;
                }
            *J.fin_ticket = 0u;  // the next launch's count starts from zero
        }
    }
    if (threadIdx.x == 0) {
        Ctl c = c0;
        if (node_open) node_phase_finish(J, c, tot[0], tot[1]);
        if (elem) elem_phase_finish(J, c, dc, tot + 2);
        *ctl = c;
    }
}
// ---------------------------------------------------------------------------
// Pipelined element pass (mixed engine, LDS staging, block assembly).
//
// A persistent grid of blocks walks the element blocks ("tiles") of the
// block map: block b takes tiles b, b + grid, b + 2*grid, ... (so a block
// stays on the XCD whose contiguous range of tiles it serves).  The next
// tile's node records travel HBM -> LDS with global_load_lds (no registers
// involved) into the second of two staging buffers while the current tile
// is computed from the first; the only exposed latency per tile is one
// round trip for its element state and the next tile's node ids.  The body
// is element_mixed_S's (same operations, same assembly order); the two
// instantiations may contract multiply-adds differently, so each tile is
// given one of the two passes by its job's properties and its own node
// count, and keeps it in any batch.  Tiles of PIPE_CAP nodes or more use
// k_element.  In jobs with amplitude velocity BCs, the tiles that hold a
// constrained node (JobDev::blk_bc) override their staged velocities in LDS
// on step ticks before the corner reads, as k_element does while staging.
//
// LDS per block is exactly the two buffers, 2 x 5 x 16 B x PIPE_CAP = 80 KiB,
// so two blocks fit a CU's 160 KiB: once every corner value of the tile has
// been read, the tile's force accumulators (4 floats per node) take the
// [X0 X1] row of the current buffer and the block reduction takes its
// [X2 1/m] row.
constexpr int PIPE_BLK = 256;
constexpr int PIPE_CAP = 512;  // staged nodes per tile; larger tiles use k_element
constexpr int PIPE_NW = PIPE_BLK / 64;                // waves per block
constexpr int PIPE_NV = (5 + PIPE_NW - 1) / PIPE_NW;  // block values reduced per wave (at most)
static_assert(PIPE_CAP == 2 * PIPE_BLK, "two staged slots per thread");
struct alignas(16) Pair {
    double a, b;
};
typedef ENSC_LDS Pair (*PipeBuf)[PIPE_CAP];  // [5][PIPE_CAP]: [X0 X1] [X2 1/m] [v0 v1] [v2 u0] [u1 u2]

// Staging of a tile's node records into a buffer: each thread owns slots
// (wave's first slot) + lane + 256*j, j < 2 (PIPE_CAP == 512).  pipe_ids
// loads the node ids of those slots; pipe_copy then issues five 16-byte
// global->LDS copies per slot, which complete in the background.
// Slots past the tile's last (and every slot of an empty range, whose n0 is
// still a valid position) read a valid id that pipe_copy does not use: the
// loads are unconditional, so no register is first zeroed and then loaded,
// which would make the compiler wait for the loads of the previous use.
__device__ __forceinline__ void pipe_ids(const JobDev& J, uint32_t n0, uint32_t nbn, uint32_t g[2]) {
    const uint32_t s0 = threadIdx.x;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const uint32_t s = s0 + j * PIPE_BLK;
        g[j] = gld(J.blk_nodes, n0 + (s < nbn ? s : (nbn ? nbn - 1 : 0u)));
    }
}
__device__ __forceinline__ void pipe_copy(const JobDev& J, uint32_t nbn, const uint32_t g[2], PipeBuf buf) {
    const uint32_t lane = threadIdx.x & 63u, w0 = threadIdx.x & ~63u;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const uint32_t s0 = w0 + j * PIPE_BLK;  // the wave's first slot: the copies land at lane * 16 from it
        if (s0 + lane < nbn) {
            const ENSC_GLB char* pp = (const ENSC_GLB char*)J.pos + g[j] * (uint32_t)(8 * POS_REC);
            const ENSC_GLB char* pv = (const ENSC_GLB char*)J.vu + g[j] * (uint32_t)(8 * VU_REC);
            __builtin_amdgcn_global_load_lds((const ENSC_GLB void*)pp, (ENSC_LDS void*)&buf[0][s0], 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const ENSC_GLB void*)(pp + 16), (ENSC_LDS void*)&buf[1][s0], 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const ENSC_GLB void*)pv, (ENSC_LDS void*)&buf[2][s0], 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const ENSC_GLB void*)(pv + 16), (ENSC_LDS void*)&buf[3][s0], 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const ENSC_GLB void*)(pv + 32), (ENSC_LDS void*)&buf[4][s0], 16, 0, 0);
        }
    }
}
// The BC override (bc_override) of the staged velocities of a tile's nodes:
// thread slots tid and tid + PIPE_BLK, each reading its node id again.
__device__ __forceinline__ void pipe_bc_patch(const JobDev& J, uint32_t n0, uint32_t nbn, double t, PipeBuf buf) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const uint32_t s = threadIdx.x + j * PIPE_BLK;
        if (s < nbn) {
            const uint32_t g = gld(J.blk_nodes, n0 + s);
            const V3 v = bc_override_inl(J, g, t, V3{buf[2][s].a, buf[2][s].b, buf[3][s].a});
            buf[2][s].a = v.x;
            buf[2][s].b = v.y;
            buf[3][s].a = v.z;
        }
    }
}

// Accumulators of the current tile in its [X0 X1] row, 4 floats per node,
// written once every wave has read its corner values (each node's first
// contribution stores, so they need no clearing).
__device__ __forceinline__ ENSC_LDS float* pipe_acc(PipeBuf buf) { return (ENSC_LDS float*)&buf[0][0]; }

// What a tile prefetches for the next one while its own corner forces and
// assembly run: the next tile's element state (into the registers the
// current state has just freed) and the node ids of the tile after that.
// Indices are clamped to valid elements and a missing tile reads nothing,
// so the loads are issued on every path.
struct PipeNext {
    const JobDev* J;   // the next tile's job
    uint32_t e;        // this thread's element of the next tile (clamped)
    const JobDev* J2;  // the job of the tile after
    uint32_t n0, nbn;  // its staging range (nbn = 0: no such tile)
    bool done;
};
__device__ __forceinline__ void pipe_prefetch(PipeNext& p, ElemState<float>& ESn, uint32_t gn[2]) {
    load_state<float, true>(*p.J, p.e, true, ESn);
    pipe_ids(*p.J2, p.n0, p.nbn, gn);
    p.done = true;
}

#define PX(n, i) ((i) == 0 ? buf[0][n].a : ((i) == 1 ? buf[0][n].b : buf[1][n].a))
#define PV(n, i) ((i) == 0 ? buf[2][n].a : ((i) == 1 ? buf[2][n].b : buf[3][n].a))
#define PU(n, i) ((i) == 0 ? buf[3][n].b : ((i) == 1 ? buf[4][n].a : buf[4][n].b))
#define NODE(k) (((k)&1) ? ((ES.nd[(k) >> 1] >> 16) & 0x7FFFu) : (ES.nd[(k) >> 1] & 0x7FFFu))
#define FIRST(k) (((k)&1) ? (ES.nd[(k) >> 1] >> 31) : ((ES.nd[(k) >> 1] >> 15) & 1u))
#define NVAL(k, i) \
    ((i) < 3 ? PX(NODE(k), (i)) + PU(NODE(k), (i)) : ((i) < 6 ? PV(NODE(k), (i)-3) : PU(NODE(k), (i)-6)))
#define CORNER_V(k, i) NVAL(k, 3 + (i))
#define ACC_W 4
// a node's first contribution adds to the zero slot's value instead of its
// stale accumulator (the tile's last staged slot, never a node of the tile:
// pipelined tiles hold fewer than PIPE_CAP nodes; zeroed at the tile start)
#define ACC_ZERO (acc + (PIPE_CAP - 1) * ACC_W)
#define ASM_BEGIN ((void)0)  // first contributions store (FIRST), nothing to clear
#define PREFETCH_HOOK pipe_prefetch(pn, ESn, gn)
__device__ __forceinline__ void element_mixed_P(const JobDev& J, uint32_t e, const ElemState<float>& ES, double dt,
                                                int32_t kind, double time_n, bool act, bool asmode, double& Le,
                                                double& dts, double* en, uint32_t& errf, const PipeBuf buf,
                                                uint32_t nbn, PipeNext& pn, ElemState<float>& ESn, uint32_t gn[2]) {
#pragma clang fp contract(fast)
    typedef float R;
    constexpr bool MIXED = true, ASM_CAP = true;
    ENSC_LDS float* acc = pipe_acc(buf);
#include "elem_body_mixed.inc"
}
#undef NODE
#undef FIRST
#undef NVAL
#undef CORNER_V
#undef PX
#undef PV
#undef PU
#undef ACC_W
#undef ACC_ZERO
#undef ASM_BEGIN
#undef PREFETCH_HOOK

// A tile's results held in registers until the next tile has started: its
// partial nodal forces (this thread's staged slots tid and tid + PIPE_BLK)
// and the block's five values (min edge, min stable dt, three work terms):
// value q in lane 63 of wave q mod PIPE_NW, as its r[q / PIPE_NW].
// Stored last in the tile, they would be the youngest memory operations at
// the next tile's start, whose wait would then also wait for their
// completion; stored right after that wait they complete during the tile.
struct PipeOut {
    float f[2][3];
    double r[PIPE_NV];
    float* pf;       // the tile's first partial record (null: nothing to store)
    double* pe;      // the tile's block partials (stride nb)
    uint32_t nbn, nb;
};
__device__ __forceinline__ void pipe_flush(PipeOut& o) {
    if (!o.pe) return;
    if (o.pf) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const uint32_t i = threadIdx.x + j * PIPE_BLK;
            if (i < o.nbn) {
                gst(o.pf, 3 * i) = o.f[j][0];
                gst(o.pf, 3 * i + 1) = o.f[j][1];
                gst(o.pf, 3 * i + 2) = o.f[j][2];
            }
        }
    }
    if ((threadIdx.x & 63) == 63) {
        const uint32_t w = threadIdx.x >> 6;
#pragma unroll
        for (int j = 0; j < PIPE_NV; ++j) {
            const uint32_t q = w + j * PIPE_NW;
            if (q < 5) gst(o.pe, q * o.nb) = o.r[j];
        }
    }
    o.pe = nullptr;
}

// One tile: compute it from `cur` (its node records) and ES (its element
// state) while the next tile's records are copied into `nxt`; g holds the
// next tile's node ids.  Fills ESn and gn for the next tile.  The one wait
// of a tile, at its start, covers this tile's copies (issued during the
// previous tile), ES and g (prefetched by the previous tile) and the
// previous tile's stores; the barrier after it also means every wave is
// done with `nxt` (the previous tile's buffer).  Returns false after the
// block's last tile.
template <bool BC>
__device__ __forceinline__ bool pipe_tile(const JobDev* __restrict__ jobs, const uint4* __restrict__ blkmap,
                                          uint32_t ntiles, uint32_t& t, PipeBuf cur, PipeBuf nxt,
                                          const ElemState<float>& ES, const uint32_t g[2], ElemState<float>& ESn,
                                          uint32_t gn[2], PipeOut& out) {
#pragma clang fp contract(off)
    const uint4 bm = blkmap[t];  // (job, block of the job, first staged position, staged count)
    const JobDev& J = jobs[bm.x];
    Ctl* ctl = J.ctl;
    const uint32_t lb = bm.y, nbn = bm.w;
    const uint32_t e = lb * PIPE_BLK + threadIdx.x;
    const bool act = e < J.ne;
    const Decision dc = decide(ctl_read(ctl), J.factor, J.wos);
    const uint32_t tn = t + gridDim.x, tnn = tn + gridDim.x;
    const bool more = tn < ntiles, more2 = tnn < ntiles;
    const uint4 bn = blkmap[more ? tn : t];
    const uint4 bnn = blkmap[more2 ? tnn : t];
    PipeNext pn;
    pn.J = &jobs[bn.x];
    {
        const uint32_t en = bn.y * PIPE_BLK + threadIdx.x;
        pn.e = en < pn.J->ne ? en : pn.J->ne - 1;
    }
    pn.J2 = &jobs[bnn.x];
    pn.n0 = bnn.z;
    pn.nbn = more2 ? bnn.w : 0u;
    pn.done = false;
    __builtin_amdgcn_s_waitcnt(0);
    lds_sync();
    pipe_flush(out);  // the previous tile's results (registers, not LDS)
    if (threadIdx.x == 0) {
        // the assembly's zero slot (ACC_ZERO), before its first phase
        ENSC_LDS float* z = pipe_acc(cur) + (PIPE_CAP - 1) * 4;
        z[0] = z[1] = z[2] = 0.0f;
    }
    if (BC && J.dyn_bc && dc.kind == KIND_STEP && gld(J.blk_bc, lb)) {
        // the velocity BCs at this step's time on the tile's constrained
        // nodes, as the staged pass applies them while staging; before the
        // next tile's copies are issued, so the wait for these loads does
        // not also wait for those (uniform branch; few tiles take it)
        pipe_bc_patch(J, bm.z, nbn, dc.time_n, cur);
        lds_sync();
    }
    if (more) pipe_copy(jobs[bn.x], bn.w, g, nxt);
    if (dc.kind != KIND_NONE) {
        const bool asmode = J.block_asm && dc.kind != KIND_PREPASS;
        double Le = __builtin_inf(), dts = __builtin_inf();
        double en[3] = {0.0, 0.0, 0.0};
        uint32_t errf = 0;
        if (asmode || act)
            element_mixed_P(J, act ? e : J.ne - 1, ES, dc.dt_i_1, dc.kind, dc.time_n, act, asmode, Le, dts, en, errf,
                            cur, nbn, pn, ESn, gn);
        if (!pn.done) pipe_prefetch(pn, ESn, gn);
        if (!act) {
            Le = dts = __builtin_inf();
            en[0] = en[1] = en[2] = 0.0;
        }
        // block partials: every thread's five values go to the [X2 1/m]
        // and [v0 v1] rows of `cur` ([5][PIPE_BLK] doubles); then wave w
        // folds the values q with q mod PIPE_NW = w over the waves' lanes in
        // wave order and reduces the wave through DPP into lane 63: one
        // wave reduction per value instead of one per value in every wave
        uint32_t werr = 0;
#pragma unroll
        for (int b = 0; b < 3; ++b)
            if (__ballot((errf >> b) & 1u)) werr |= 1u << b;
        if (werr && (threadIdx.x & 63) == 0) atomicOr(&ctl->err, werr);
        ENSC_LDS double(*s_val)[PIPE_BLK] = (ENSC_LDS double(*)[PIPE_BLK]) & cur[1][0];
        auto put_red = [&]() {
            const uint32_t i = threadIdx.x;
            s_val[0][i] = Le;
            s_val[1][i] = dts;
            s_val[2][i] = en[0];
            s_val[3][i] = en[1];
            s_val[4][i] = en[2];
        };
        out.pf = nullptr;
        if (asmode) {
            // the rows are free: every wave is past the assembly's barriers,
            // the first of which follows every corner read; the barrier below
            // ends the last assembly phase and publishes the values
            put_red();
            const ENSC_LDS float* acc = pipe_acc(cur);
            lds_sync();
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const uint32_t i = threadIdx.x + j * PIPE_BLK;
                const uint32_t k = i < nbn ? i : 0u;
                out.f[j][0] = acc[4 * k];
                out.f[j][1] = acc[4 * k + 1];
                out.f[j][2] = acc[4 * k + 2];
            }
            out.pf = static_cast<float*>(J.part_f) + (size_t)bm.z * 3;
            out.nbn = nbn;
        } else {
            lds_sync();  // every wave's corner reads are done
            put_red();
            lds_sync();
        }
        {
            const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;  // w is uniform in the wave
            auto fold = [&](uint32_t q, bool mn) {
                double r = s_val[q][lane];
#pragma unroll
                for (int x = 1; x < PIPE_NW; ++x) {
                    const double v = s_val[q][lane + 64 * x];
                    r = mn ? (v < r ? v : r) : r + v;
                }
                return mn ? wave_reduce63<true>(r) : wave_reduce63<false>(r);
            };
#pragma unroll
            for (int j = 0; j < PIPE_NV; ++j) {
                const uint32_t q = w + j * PIPE_NW;
                if (q < 2)
                    out.r[j] = fold(q, true);  // min edge, min stable dt
                else if (q < 5)
                    out.r[j] = fold(q, false);  // internal, hourglass and viscous work
            }
        }
        out.pe = J.part_e + lb;
        out.nb = J.nblk_e;
    } else {
        pipe_prefetch(pn, ESn, gn);
    }
    if (lb == 0 && threadIdx.x == 0) gst(&ctl->tick_elem, 0u) = dc.kind;  // the tick record
    t = tn;
    return more;
}

// BC: the launch holds tiles of jobs with amplitude velocity BCs (the
// override of constrained tiles' velocities is compiled in; without it the
// tile loop is leaner).  Both instantiations run the same arithmetic.
template <int WPE, bool BC>
__global__ __launch_bounds__(PIPE_BLK) __attribute__((amdgpu_waves_per_eu(WPE))) void k_element_pipe(
    const JobDev* __restrict__ jobs, const uint4* __restrict__ blkmap, uint32_t ntiles) {
    __shared__ Pair s_bufA[5][PIPE_CAP];
    __shared__ Pair s_bufB[5][PIPE_CAP];
    const PipeBuf bufA = (PipeBuf)s_bufA;
    const PipeBuf bufB = (PipeBuf)s_bufB;
    uint32_t t = blockIdx.x;
    if (t >= ntiles) return;
    ElemState<float> esA, esB;
    uint32_t gA[2], gB[2];
    {
        // the first tile's copies and state, the second tile's node ids
        const uint4 bm = blkmap[t];
        const JobDev& J = jobs[bm.x];
        pipe_ids(J, bm.z, bm.w, gA);
        pipe_copy(J, bm.w, gA, bufA);
        const uint32_t e = bm.y * PIPE_BLK + threadIdx.x;
        load_state<float, true>(J, e < J.ne ? e : J.ne - 1, true, esA);
        const uint32_t tn = t + gridDim.x;
        const uint4 bn = blkmap[tn < ntiles ? tn : t];
        pipe_ids(jobs[bn.x], bn.z, tn < ntiles ? bn.w : 0u, gB);
    }
    // two tiles per trip so each buffer is a distinct LDS object in every
    // access (the compiler then keeps copies into one buffer in flight while
    // the other is read)
    PipeOut out;
    out.pe = nullptr;
    while (pipe_tile<BC>(jobs, blkmap, ntiles, t, bufA, bufB, esA, gB, esB, gA, out) &&
           pipe_tile<BC>(jobs, blkmap, ntiles, t, bufB, bufA, esB, gA, esA, gB, out)) {
    }
    pipe_flush(out);  // the block's last tile
}

// ASM: the launch covers block-assembly jobs only (mixed engine) and sums
// their block partials; otherwise the per-corner forces through the ELL.
template <bool MIXED, int BLK, bool ASM>
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(ASM ? (MIXED ? 7 : 4) : 1))) void k_node(
    const JobDev* __restrict__ jobs, const uint4* __restrict__ blkmap) {
#pragma clang fp contract(off)
    const uint4 bm = blkmap[blockIdx.x];  // (job, block of the job, -, -)
    const JobDev& J = jobs[bm.x];
    Ctl* ctl = J.ctl;
    const Ctl cs = ctl_read(ctl);
    if (cs.kind != KIND_STEP) {
        if (bm.y == 0 && threadIdx.x == 0) gst(&ctl->tick_nsteps, 0u) = 0;  // the tick record
        return;
    }
    const double dt_i = cs.dt_i, dt_i_1 = cs.dt_i_1, time = cs.time;
    const int32_t snap = cs.snap_now - 1;  // output slot of this step, or -1
    const uint32_t lb = bm.y;
    const uint32_t n = lb * BLK + threadIdx.x;
    double ke = 0.0, pw = 0.0;
    if (snap >= 0 && n < J.nn) {
        // output_results' element stress (this step's, already updated)
        typedef typename std::conditional<MIXED, float, double>::type R_;
        // global (not generic) accesses: generic ones pending here would make
        // the compiler's later waits conservative on every path
        ENSC_GLB double* se = (ENSC_GLB double*)(J.snap_elem + (size_t)snap * J.ne * 6);
        const ENSC_GLB R_* sg = (const ENSC_GLB R_*)J.sigma;
        for (uint32_t e = n; e < J.ne; e += J.nn)
#pragma unroll
            for (int q = 0; q < 6; ++q)
                se[(size_t)e * 6 + q] = (double)sg[(size_t)q * J.ne_s + e];
    }
    if (n < J.nn) {
        // [X0 X1 | X2 im] [v0 v1 | v2 u0 | u1 u2]; the mixed engine never
        // rewrites X (the reference position), so it reads 1/m from its own
        // array instead of the position record
        // block assembly: the node's range of partials first, so the loads
        // that depend on it wait for it alone
        uint32_t q0 = 0, q1 = 0;
        if (ASM) {
            q0 = gld(J.pos_ptr, n);
            q1 = gld(J.pos_ptr, n + 1);
        }
        const D2 p2 = gld2(J.vu, n * VU_REC), p3 = gld2(J.vu, n * VU_REC + 2), p4 = gld2(J.vu, n * VU_REC + 4);
        double x[3] = {0.0, 0.0, 0.0};
        double im;
        if (MIXED) {
            im = gld(J.inv_m, n);
        } else {
            const D2 p0 = gld2(J.pos, n * POS_REC), p1 = gld2(J.pos, n * POS_REC + 2);
            x[0] = p0.a;
            x[1] = p0.b;
            x[2] = p1.a;
            im = p1.b;
        }
        double v[3] = {p2.a, p2.b, p3.a};
        double u[3] = {p3.b, p4.a, p4.b};
        double acc[3] = {0.0, 0.0, 0.0};
        if (ASM) {
            // block assembly: the node's partials, one per element block that
            // holds it, in ascending block order
            typedef typename std::conditional<MIXED, float, double>::type P;
            for (uint32_t q = q0; q < q1; q += 4) {
                // four positions, then four records, each set of loads issued
                // together (indices past the node's last clamped to it; their
                // values are weighted by 0, so no load depends on a branch)
                uint32_t pos[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) pos[j] = gld(J.pos_list, q + j < q1 ? q + j : q1 - 1);
                P f[4][3];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    // one record: a single address, three offsets
                    const ENSC_GLB P* pf =
                        (const ENSC_GLB P*)((const ENSC_GLB char*)J.part_f + pos[j] * (uint32_t)(3 * sizeof(P)));
                    f[j][0] = pf[0];
                    f[j][1] = pf[1];
                    f[j][2] = pf[2];
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const double w = (q + j < q1) ? 1.0 : 0.0;
                    acc[0] += (double)f[j][0] * w;
                    acc[1] += (double)f[j][1] * w;
                    acc[2] += (double)f[j][2] * w;
                }
            }
        } else {
        // Element forces of the node in (corner, original element) order;
        // slots are filled from the front, ELL_EMPTY after the last.  Each
        // chunk of 8 slots issues all its loads before summing.
        double grp[3] = {0.0, 0.0, 0.0};
        int gk = -1;
        const uint32_t es = J.ne_s, W = (uint32_t)J.ell_w;
        for (uint32_t w0 = 0; w0 < W; w0 += 8) {
            uint32_t sl[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) sl[j] = (w0 + j < W) ? gld(J.ell, (w0 + j) * J.nn_s + n) : ELL_EMPTY;
            double f[8][3];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                f[j][0] = f[j][1] = f[j][2] = 0.0;
                if (sl[j] != ELL_EMPTY) {
                    const uint32_t idx = (sl[j] & 7u) * es + (sl[j] >> 3);
                    if (MIXED) {
                        const float* q = static_cast<const float*>(J.fel);
                        f[j][0] = gld(q, 3 * idx);
                        f[j][1] = gld(q, 3 * idx + 1);
                        f[j][2] = gld(q, 3 * idx + 2);
                    } else {
                        const double* q = static_cast<const double*>(J.fel);
                        f[j][0] = gld(q, 3 * idx);
                        f[j][1] = gld(q, 3 * idx + 1);
                        f[j][2] = gld(q, 3 * idx + 2);
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                if (sl[j] == ELL_EMPTY) continue;
                const int k = (int)(sl[j] & 7u);
                if (k != gk) {
                    if (gk >= 0) {
                        acc[0] += grp[0];
                        acc[1] += grp[1];
                        acc[2] += grp[2];
                    }
                    grp[0] = grp[1] = grp[2] = 0.0;
                    gk = k;
                }
                grp[0] += f[j][0];
                grp[1] += f[j][1];
                grp[2] += f[j][2];
            }
        }
        if (gk >= 0) {
            acc[0] += grp[0];
            acc[1] += grp[1];
            acc[2] += grp[2];
        }
        }
        // apply_fsiSph_nodeForce (1635-1689) runs before add_inForce_to_rhs:
        // a = (sum of the particles' nodal forces) + (the corner sum)
        int32_t mf = -1;
        if (J.n_fsi_node) mf = gld(J.fsi_node_of, n);
        double a[3];
        if (mf >= 0) {
#pragma unroll
            for (int d = 0; d < 3; ++d) a[d] = gld(J.fsi_nforce, (uint32_t)mf * 3 + d) + acc[d];
        } else {
#pragma unroll
            for (int d = 0; d < 3; ++d) a[d] = 0.0 + acc[d];
        }
        if (J.grav_on) {  // apply_gravity (800-826)
            const double ag = cs.grav_amp * J.grav_g;
            a[0] += ag * J.grav_dir[0] / im;
            a[1] += ag * J.grav_dir[1] / im;
            a[2] += ag * J.grav_dir[2] / im;
        }
        // surface pressures (extension): the forces of the node's loaded faces,
        // in ascending face order; their work is counted like gravity's (the
        // update loop sums the forces again, so nothing stays live between)
        uint32_t f0 = 0, f1 = 0;
        if (J.n_dsf) {
            f0 = gld(J.ds_ptr, n);
            f1 = gld(J.ds_ptr, n + 1);
            for (uint32_t q = f0; q < f1; ++q) {
                const uint32_t fq = gld(J.ds_list, q) * 3;
#pragma unroll
                for (int d = 0; d < 3; ++d) a[d] += gld(J.dsf_force, fq + d);
            }
        }
        double wp = 0.0;
        // External work of the step (extension; the reference keeps no energy
        // account), per node:
        //   wr: the reactions' power at velocity-BC dofs, -(f_int + f_grav) v_c,
        //       v_c the constrained velocity (the BC value, or 0 where an SPC
        //       line also holds the dof);
        //   wg: gravity's power on every dof over g m, dir . (v_old + v_new),
        //       v_old the velocity after the BCs;
        //   wj: twice the kinetic energy over m that a constraint adds or
        //       removes when it sets v (a ramped velocity BC), v_c^2 - v_record^2.
        // Work = dt_avg (wr + ag m wg / 2) + m wj / 2, dt_avg = (dt_i + dt_i_1) / 2;
        // the constraint terms are folded into pw where they are formed.
        const double hs = dt_i + dt_i_1;
        double wg = 0.0, kj = 0.0;
        const uint8_t mk = gld(J.bcmask, n);
        double vbc[3] = {0.0, 0.0, 0.0};  // constrained velocity of each constrained dof
        if (mk) {  // apply_boundary_condition_a (2029-2145): velocity lines, then SPC lines
            double rfv[3] = {0.0, 0.0, 0.0};
            double wr = 0.0, wj = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                if (mk & (BC_VEL0 << d)) {
                    vbc[d] = gld(J.bcval, n * 3 + d) * amp_value(J, gld(J.bcamp, n * 3 + d), time);
                    if (!(mk & (BC_SPC0 << d))) wr += -a[d] * vbc[d];
                    a[d] = 0.0;
                }
                if (mk & (BC_SPC0 << d)) {
                    vbc[d] = 0.0;
                    if (a[d] != 0.0) {
                        rfv[d] = -a[d];
                        a[d] = 0.0;
                    }
                }
                if (mk & ((BC_VEL0 | BC_SPC0) << d)) wj += vbc[d] * vbc[d] - v[d] * v[d];
            }
            pw = 0.5 * hs * wr;
            kj = wj;
            if (mk & (7u * BC_SPC0)) {
                gst(J.rf, n * 3) = rfv[0];
                gst(J.rf, n * 3 + 1) = rfv[1];
                gst(J.rf, n * 3 + 2) = rfv[2];
            }
        }
        double k2 = 0.0;
        const uint32_t cw = MIXED ? gld(J.ucomp, n) : 0u;
        uint32_t cw_new = 0;
        ENSC_GLB double* sn = snap >= 0 ? (ENSC_GLB double*)(J.snap_node + ((size_t)snap * J.nn + n) * 12) : nullptr;
        if (sn) {
            // output_results' node state: x and u before this step's update
            double xs[3] = {x[0], x[1], x[2]};
            if (MIXED) {
                const D2 p0 = gld2(J.pos, n * POS_REC), p1 = gld2(J.pos, n * POS_REC + 2);
                xs[0] = p0.a + u[0];
                xs[1] = p0.b + u[1];
                xs[2] = p1.a + u[2];
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                sn[d] = xs[d];
                sn[9 + d] = u[d];
            }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double ad = a[d] * im;  // compute_a (642-644)
            double vd = v[d];
            if (mk & ((BC_VEL0 | BC_SPC0) << d)) vd = vbc[d];
            if (sn) {
                // a after compute_a; v after the velocity BCs
                sn[3 + d] = ad;
                sn[6 + d] = vd;
            }
            const double vb = vd;
            vd += hs * ad / 2.0;  // update_velocity (646-649)
            wg += J.grav_dir[d] * (vb + vd);
            if (f1 != f0) {
                double fp = 0.0;
                for (uint32_t q = f0; q < f1; ++q) fp += gld(J.dsf_force, gld(J.ds_list, q) * 3 + d);
                wp += fp * (vb + vd);
            }
            if (mf >= 0) wp += gld(J.fsi_nforce, (uint32_t)mf * 3 + d) * (vb + vd);  // the particles' work
            const double dx = dt_i * vd;
            if (MIXED) {
                // update_displacement (651-654), compensated (Kahan); the
                // position is X + u, so move_mesh has nothing to store
                const double y = dx - ucomp_get(cw, d, u[d]);
                const double t = u[d] + y;
                cw_new |= ucomp_put((t - u[d]) - y, t, d);
                u[d] = t;
            } else {
                u[d] += dx;  // update_displacement (651-654)
                x[d] += dx;  // move_mesh (688-696)
            }
            v[d] = vd;
            k2 += vd * vd;
        }
        // kinetic energy, and the kinetic-energy change of the constrained
        // dofs, with one division by 1/m
        const double half_m = 0.5 / im;
        ke = k2 * half_m;
        pw += kj * half_m;
        if (J.grav_on) pw += 0.5 * hs * (cs.grav_amp * J.grav_g) * wg * half_m;
        if (f1 != f0 || mf >= 0) pw += 0.25 * hs * wp;
        if (MIXED) gst(J.ucomp, n) = cw_new;
        if (!MIXED) {
            gst2(J.pos, n * POS_REC, x[0], x[1]);
            gst2(J.pos, n * POS_REC + 2, x[2], im);
        }
        gst2(J.vu, n * VU_REC, v[0], v[1]);
        gst2(J.vu, n * VU_REC + 2, v[2], u[0]);
        gst2(J.vu, n * VU_REC + 4, u[1], u[2]);
    }
    __shared__ double s_red[BLK / 64][2];
    ke = wave_reduce63<false>(ke);
    pw = wave_reduce63<false>(pw);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) {
        s_red[w][0] = ke;
        s_red[w][1] = pw;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r0 = s_red[0][0], r1 = s_red[0][1];
        for (int q = 1; q < BLK / 64; ++q) {
            r0 += s_red[q][0];
            r1 += s_red[q][1];
        }
        double* p = J.part_n + lb;
        p[0] = r0;
        p[J.nblk_n] = r1;
        if (lb == 0) gst(&ctl->tick_nsteps, 0u) = cs.nsteps;  // the tick record: an update is open
    }
}

// Surface pressures (extension; *Dsload): one thread per loaded face of the
// block's job computes the force its pressure puts on each of its 4 nodes
// this step, p amp A / 4 with A = (x2 - x0) x (x3 - x1) / 2 the face's vector
// area (into the element) on the positions at the start of the step, and the
// amplitude averaged over the step as for gravity (deal_amp).  Runs between
// the element finalize (which sets the step's time and dt) and the nodal
// update, which adds the forces of each node's faces.
constexpr int FACE_BLK = 256;
__global__ __launch_bounds__(FACE_BLK) void k_face(const JobDev* __restrict__ jobs, const uint4* __restrict__ blkmap) {
#pragma clang fp contract(off)
    const uint4 bm = blkmap[blockIdx.x];  // (job, block of the job's faces, -, -)
    const JobDev& J = jobs[bm.x];
    const Ctl* ctl = J.ctl;
    if (ctl->kind != KIND_STEP) return;
    const uint32_t f = bm.y * FACE_BLK + threadIdx.x;
    if (f >= J.n_dsf) return;
    const bool moving_x = J.inv_m == nullptr;  // FP64 jobs store the current position; mixed ones X (+ u)
    double x[4][3];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const uint32_t n = J.dsf_nodes[4 * f + c];
#pragma unroll
        for (int d = 0; d < 3; ++d)
            x[c][d] = J.pos[(size_t)n * POS_REC + d] + (moving_x ? 0.0 : J.vu[(size_t)n * VU_REC + 3 + d]);
    }
    const double a[3] = {x[2][0] - x[0][0], x[2][1] - x[0][1], x[2][2] - x[0][2]};
    const double b[3] = {x[3][0] - x[1][0], x[3][1] - x[1][1], x[3][2] - x[1][2]};
    const double amp = J.dsf_amp[f] >= 0 ? deal_amp(J, J.dsf_amp[f], ctl->time, ctl->dt_i) : 1.0;
    const double s = J.dsf_val[f] * amp * 0.125;  // p amp / 4 times the 1/2 of the cross product
    double* out = J.dsf_force + (size_t)f * 3;
    out[0] = s * (a[1] * b[2] - a[2] * b[1]);
    out[1] = s * (a[2] * b[0] - a[0] * b[2]);
    out[2] = s * (a[0] * b[1] - a[1] * b[0]);
}

// Trilinear shape function of corner k at unit point p
// (Element_HexN8::get_shapeFunctionValue, Element_HexN8.cpp:46-93: corner k
// at (-1|1, -1|1, -1|1) in the order 0 (-,-,-), 1 (+,-,-), 2 (+,+,-),
// 3 (-,+,-), 4..7 the same at +1 in the third coordinate).
__device__ inline double hex_shape(int k, const double* p) {
#pragma clang fp contract(off)
    const double c0 = ((k + 1) & 2) ? 1.0 : -1.0;
    const double c1 = (k & 2) ? 1.0 : -1.0;
    const double c2 = (k & 4) ? 1.0 : -1.0;
    return 1.0 / 8.0 * (1 + c0 * p[0]) * (1 + c1 * p[1]) * (1 + c2 * p[2]);
}

// Virtual-particle forces onto the nodes (exDyna3D::apply_fsiSph_nodeForce,
// exDyna3D.cpp:1635-1689): one thread per FSI node sums, over its
// (element, corner) entries in ascending element order, the element's
// nodal force at that corner, sum over the element's particles of
// f_p N_corner(unit_p) in particle order.  The nodal update adds the result
// before the internal forces.  Runs on the single job of the engine (job 0).
constexpr int FSI_BLK = 256;
__global__ __launch_bounds__(FSI_BLK) void k_fsi_force(const JobDev* __restrict__ jobs) {
#pragma clang fp contract(off)
    const JobDev& J = jobs[0];
    if (J.ctl->kind != KIND_STEP) return;
    const uint32_t m = blockIdx.x * FSI_BLK + threadIdx.x;
    if (m >= J.n_fsi_node) return;
    double f[3] = {0.0, 0.0, 0.0};
    for (uint32_t q = J.fsi_nptr[m]; q < J.fsi_nptr[m + 1]; ++q) {
        const uint32_t ent = J.fsi_nent[q], s = ent >> 3;
        const int k = (int)(ent & 7u);
        double nf[3] = {0.0, 0.0, 0.0};
        for (uint32_t p = J.fsi_eptr[s]; p < J.fsi_eptr[s + 1]; ++p) {
            const double N = hex_shape(k, J.fsi_unit + (size_t)p * 3);
#pragma unroll
            for (int d = 0; d < 3; ++d) nf[d] += J.fsi_force[(size_t)p * 3 + d] * N;
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) f[d] += nf[d];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) J.fsi_nforce[(size_t)m * 3 + d] = f[d];
}

// Virtual-particle positions and velocities from the element's nodes
// (exDyna3D::update_virParticles_coor_vel, exDyna3D.cpp:1691-1758, and the
// generation in generateVirtualParticlesForFace, 1428-1521): x_p = sum_j
// N_j(unit_p) x_j, v_p likewise, after the nodal update.  One thread per
// particle of job 0.
__global__ __launch_bounds__(FSI_BLK) void k_fsi_particles(const JobDev* __restrict__ jobs) {
#pragma clang fp contract(off)
    const JobDev& J = jobs[0];
    const uint32_t p = blockIdx.x * FSI_BLK + threadIdx.x;
    if (p >= J.n_fsi_part) return;
    const uint32_t* en = J.fsi_econn + (size_t)J.fsi_pelem[p] * 8;
    const double* up = J.fsi_unit + (size_t)p * 3;
    const bool moving_x = J.inv_m == nullptr;  // FP64 jobs store the current position; mixed ones X (+ u)
    double x[3], v[3];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const double N = hex_shape(j, up);
        const size_t n = en[j];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double xj = J.pos[n * POS_REC + d] + (moving_x ? 0.0 : J.vu[n * VU_REC + 3 + d]);
            const double vj = J.vu[n * VU_REC + d];
            x[d] = j ? x[d] + N * xj : N * xj;
            v[d] = j ? v[d] + N * vj : N * vj;
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        J.fsi_coord[(size_t)p * 3 + d] = x[d];
        J.fsi_vel[(size_t)p * 3 + d] = v[d];
    }
}

// Per-job node finalize (one block per job in the job table): closes the
// last nodal update of a run of step iterations.  The element finalize of
// the next iteration does the same work first, so only the last iteration
// of a run launches it.
template <int BLK>
__global__ __launch_bounds__(BLK) void k_fin_node(const JobDev* __restrict__ jobs) {
    const JobDev& J = jobs[blockIdx.x];
    if (J.ctl->kind != KIND_STEP) return;
    double acc[7];
    fold_partials<BLK, 0x03u>(J, acc);
    __shared__ double s_red[BLK / 64][7];
    double tot[7];
    combine_partials<BLK, 0x03u>(acc, s_red, tot);
    // once per run of ticks: the device copy is updated in place
    if (threadIdx.x == 0) node_phase_finish(J, *J.ctl, tot[0], tot[1]);
}

// setCurrentStep's apply_boundary_condition_vec (exDyna3D.cpp:1911, 379-490)
// at the current time, plus stepEndTime = time + timePeriod (83).  One block
// walks the listed BC nodes.  The kinetic energy the constraints add or
// remove by setting v (e.g. an initial velocity on nodes a step then fixes)
// is external work: it is summed in a fixed order (per thread, then a fixed
// tree) and added to the job's W_ext, so the energy balance
// KE + IE + HG + VIS - W_ext = KE0 also holds across step switches.
constexpr int APPLY_BC_BLK = 256;
__global__ __launch_bounds__(APPLY_BC_BLK) void k_apply_bc(JobDev* __restrict__ jobs, uint32_t job,
                                                           const uint32_t* __restrict__ nodes, uint32_t n,
                                                           double period, int set_end) {
#pragma clang fp contract(off)
    JobDev& J = jobs[job];
    const double time = J.ctl->time;
    double wj = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += APPLY_BC_BLK) {
        const uint32_t nd = nodes[i];
        const uint8_t mk = J.bcmask[nd];
        double* rec = J.vu + (size_t)nd * VU_REC;
        double w = 0.0;
        for (int d = 0; d < 3; ++d) {
            const double v0 = rec[d];
            double v1 = v0;
            if (mk & (BC_VEL0 << d)) v1 = J.bcval[(size_t)nd * 3 + d] * amp_value(J, J.bcamp[(size_t)nd * 3 + d], time);
            if (mk & (BC_SPC0 << d)) v1 = 0.0;
            rec[d] = v1;
            w += v1 * v1 - v0 * v0;
        }
        wj += 0.5 * w / J.pos[(size_t)nd * POS_REC + 3];
    }
    __shared__ double s_w[APPLY_BC_BLK / 64];
    wj = wave_sum(wj);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = wj;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = s_w[0];
        for (int q = 1; q < APPLY_BC_BLK / 64; ++q) t += s_w[q];
        J.ctl->wext += t;
        if (set_end) J.ctl->step_end = time + period;
    }
}

// Fields output_results writes for a captured snapshot (exDyna3D.cpp:
// 228-309, computeVonMises 1111-1133), one element per thread, FP64:
//   out[0]     von Mises stress of the element's stress;
//   out[1..6]  the stress in the order handed to VTK: xx yy zz xy yz xz;
//   out[7..12] Green-Lagrange strain 0.5 (F^T F - I), same order, with
//              F = I + u B0 from the snapshot's displacements (the F the
//              step's element pass formed, exDyna3D.cpp:864).
// The FP64 engine rebuilds B0 from its stored Gi in the element body's
// order; the mixed engine, which stores no Gi, forms Gi in FP64 from the
// reference positions (Gi = (XL^-1)^T, XL[j][m] = sum_k X_k[j] c_k[m]).
template <bool MIXED>
__global__ void k_post(const JobDev* __restrict__ job, int32_t slot, double* __restrict__ out) {
#pragma clang fp contract(off)
    const JobDev& J = *job;
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= J.ne) return;
    const double* sn = J.snap_node + (size_t)slot * J.nn * 12;
    const double* se = J.snap_elem + ((size_t)slot * J.ne + e) * 6;
    uint32_t nd[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) nd[k] = J.conn[(size_t)k * J.ne_s + e];
    double G[3][3];
    if (MIXED) {
        double XL[3][3];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                double t = 0.0;
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const double X = J.pos[(size_t)nd[k] * POS_REC + j];
                    t = CS_[k][m] > 0 ? t + X : t - X;
                }
                XL[j][m] = t;
            }
        double XLi[3][3];
        inverse3<double>(XL, XLi);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) G[i][j] = XLi[j][i];
    } else {
        const double* Gp = static_cast<const double*>(J.Gi);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) G[i][j] = Gp[(size_t)(3 * i + j) * J.ne_s + e];
    }
    double F[3][3];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        double Bk[3];
        bgrad<double>(G, k, Bk);
        const double* uk = sn + (size_t)nd[k] * 12 + 9;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) F[i][j] = (k == 0) ? uk[i] * Bk[j] : F[i][j] + uk[i] * Bk[j];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) F[i][j] = (i == j ? 1.0 : 0.0) + F[i][j];
    double E[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double ftf = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
            E[i][j] = 0.5 * (ftf - (i == j ? 1.0 : 0.0));
        }
    const double t1 = (se[0] - se[1]) * (se[0] - se[1]);
    const double t2 = (se[1] - se[2]) * (se[1] - se[2]);
    const double t3 = (se[2] - se[0]) * (se[2] - se[0]);
    const double t4 = 6.0 * (se[3] * se[3] + se[4] * se[4] + se[5] * se[5]);
    double* o = out + (size_t)e * 13;
    o[0] = sqrt((t1 + t2 + t3 + t4) / 2.0);
    o[1] = se[0];
    o[2] = se[1];
    o[3] = se[2];
    o[4] = se[3];
    o[5] = se[5];
    o[6] = se[4];
    o[7] = E[0][0];
    o[8] = E[1][1];
    o[9] = E[2][2];
    o[10] = E[0][1];
    o[11] = E[1][2];
    o[12] = E[0][2];
}

}  // namespace ensc
