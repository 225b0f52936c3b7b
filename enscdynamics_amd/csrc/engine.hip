// Host side of the step engine and its C ABI (include/ensc_engine.h).
//
// Replaces, for the hot path, the reference's exDyna3D::init / run /
// setCurrentStep / computeSate (Explicit3D/source/exDyna3D.cpp:54-205,
// 492-515, 642-696, 774-793, 1760-1940).  Setup quantities (B0, V0, lumped
// mass) are computed once on the host in FP64 with the reference's
// arithmetic; everything per step runs on the device (kernels.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "engine_host.h"

namespace ensc {

thread_local std::string g_last_error;

ensc_status fail(ensc_status s, const std::string& msg) {
    g_last_error = msg;
    return s;
}

#define HIPCHK(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return fail(ENSC_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------------------
// setup arithmetic (FP64, reference order)
// ---------------------------------------------------------------------------

static double det3h(const double m[3][3]) {
    const double h0 = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]);
    const double h1 = m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]);
    const double h2 = m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    return h0 - h1 + h2;
}
static double cofh(const double m[3][3], int i, int j) {
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return m[i1][j1] * m[i2][j2] - m[i1][j2] * m[i2][j1];
}
static void inv3h(const double m[3][3], double r[3][3]) {
    const double c00 = cofh(m, 0, 0), c10 = cofh(m, 1, 0), c20 = cofh(m, 2, 0);
    const double det = c00 * m[0][0] + c10 * m[1][0] + c20 * m[2][0];
    const double invdet = 1.0 / det;
    r[1][0] = cofh(m, 0, 1) * invdet;
    r[1][1] = cofh(m, 1, 1) * invdet;
    r[2][0] = cofh(m, 0, 2) * invdet;
    r[1][2] = cofh(m, 2, 1) * invdet;
    r[2][1] = cofh(m, 1, 2) * invdet;
    r[2][2] = cofh(m, 2, 2) * invdet;
    r[0][0] = c00 * invdet;
    r[0][1] = c10 * invdet;
    r[0][2] = c20 * invdet;
}

// FEValues<8,3>::reinit at the single Gauss point (fevalues.h:233-265) and
// compute_invMassMatrix (exDyna3D.cpp:492-515), original element order.
//
// B0 = Ji * dN with every dN entry +-1/8, so B0(k,i) = sum_j G(i,j) * sign(k,j)
// for G = Ji / 8; multiplying by 1/8 is exact, and the kernels form that sum
// in the reference's order (j = 0, 1, 2), so storing G (9 values) instead of
// B0 (24 values) reproduces B0 bit for bit.
static bool setup_geometry(const std::vector<double>& x0, const std::vector<uint32_t>& conn, size_t Ne, double rho,
                           std::vector<double>& G8, std::vector<double>& V0, std::vector<double>& invM, size_t Nn,
                           std::string& err) {
    static const double cs[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                    {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    double dN[3][8];
    for (int a = 0; a < 8; ++a)
        for (int d = 0; d < 3; ++d) dN[d][a] = 1.0 / 8.0 * cs[a][d];
    G8.assign(9 * Ne, 0.0);
    V0.assign(Ne, 0.0);
    std::vector<double> mass(Nn, 0.0);
    for (size_t e = 0; e < Ne; ++e) {
        double X[8][3];
        for (int k = 0; k < 8; ++k)
            for (int d = 0; d < 3; ++d) X[k][d] = x0[3 * (size_t)conn[8 * e + k] + d];
        double Jm[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s = dN[i][0] * X[0][j];
                for (int k = 1; k < 8; ++k) s += dN[i][k] * X[k][j];
                Jm[i][j] = s;
            }
        double Ji[3][3];
        inv3h(Jm, Ji);
        const double det = det3h(Jm);
        V0[e] = det * 8.0;
        if (!(V0[e] > 0.0)) {
            err = "element " + std::to_string(e) + " has non-positive reference volume (inverted or degenerate)";
            return false;
        }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) G8[9 * e + 3 * i + j] = Ji[i][j] * 0.125;
        const double nodeMass = rho * V0[e] / 8.0;
        for (int k = 0; k < 8; ++k) mass[conn[8 * e + k]] += nodeMass;
    }
    invM.resize(Nn);
    for (size_t i = 0; i < Nn; ++i) invM[i] = 1.0 / mass[i];
    return true;
}

// Space-filling (Morton) element order from element centroids, quantized
// with one scale for all three axes, so that runs of consecutive elements
// are compact in space for any aspect ratio; nodes are then numbered in
// first-touch order of the reordered elements.  The quantization cell is
// the mean element size (the cube root of the bounding box volume per
// element): on a structured mesh each element then has its own integer
// cell, the Z-curve's cells line up with the element grid, and a run of 256
// elements is an 8 x 8 x 4 box (405 nodes); a finer scale, unaligned with
// the grid, cuts element columns and makes the runs ragged (the 100^3 cube:
// 444 nodes per run on average; a 200 x 40 x 40 block: 470, some over 512).
static void morton_order(const std::vector<double>& x0, const std::vector<uint32_t>& conn, size_t Ne, size_t Nn,
                         std::vector<uint32_t>& e_new2old, std::vector<uint32_t>& n_new2old) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (size_t i = 0; i < Nn; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], x0[3 * i + d]);
            hi[d] = std::max(hi[d], x0[3 * i + d]);
        }
    auto spread = [](uint64_t v) {
        v &= 0x1fffff;
        v = (v | v << 32) & 0x1f00000000ffffULL;
        v = (v | v << 16) & 0x1f0000ff0000ffULL;
        v = (v | v << 8) & 0x100f00f00f00f00fULL;
        v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
        v = (v | v << 2) & 0x1249249249249249ULL;
        return v;
    };
    const double span = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    double h = std::cbrt((hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]) / (double)std::max<size_t>(Ne, 1));
    if (!(h > 0.0) || !std::isfinite(h)) h = 0.0;
    // at most 2^21 cells along the longest extent (21 bits per axis)
    const double cell = std::max(h, span / 2097151.0);
    std::vector<std::pair<uint64_t, uint32_t>> key(Ne);
    for (size_t e = 0; e < Ne; ++e) {
        double c[3] = {0, 0, 0};
        for (int k = 0; k < 8; ++k)
            for (int d = 0; d < 3; ++d) c[d] += x0[3 * (size_t)conn[8 * e + k] + d];
        uint64_t q[3];
        for (int d = 0; d < 3; ++d) {
            const double t = cell > 0.0 ? std::floor((c[d] / 8.0 - lo[d]) / cell) : 0.0;
            q[d] = (uint64_t)std::min(2097151.0, std::max(0.0, t));
        }
        key[e] = {spread(q[0]) | spread(q[1]) << 1 | spread(q[2]) << 2, (uint32_t)e};
    }
    std::sort(key.begin(), key.end());
    e_new2old.resize(Ne);
    for (size_t i = 0; i < Ne; ++i) e_new2old[i] = key[i].second;
    std::vector<uint32_t> seen(Nn, 0xFFFFFFFFu);
    n_new2old.clear();
    n_new2old.reserve(Nn);
    for (size_t i = 0; i < Ne; ++i)
        for (int k = 0; k < 8; ++k) {
            const uint32_t n = conn[8 * (size_t)e_new2old[i] + k];
            if (seen[n] == 0xFFFFFFFFu) {
                seen[n] = 1;
                n_new2old.push_back(n);
            }
        }
    for (size_t n = 0; n < Nn; ++n)
        if (seen[n] == 0xFFFFFFFFu) n_new2old.push_back((uint32_t)n);  // nodes in no element
}

// ---------------------------------------------------------------------------
// model copy + per-job device state
// ---------------------------------------------------------------------------

size_t round_up(size_t v, size_t m) { return (v + m - 1) / m * m; }

bool copy_model(const ensc_model_desc* m, ModelCopy& c, std::string& err) {
    if (!m || !m->x0 || !m->conn || m->n_nodes == 0 || m->n_elems == 0) {
        err = "model has no nodes or elements";
        return false;
    }
    if (m->n_steps < 2 || !m->steps) {
        err = "model needs the Initial step and at least one *Step";
        return false;
    }
    c.Nn = m->n_nodes;
    c.Ne = m->n_elems;
    c.x0.assign(m->x0, m->x0 + 3 * (size_t)c.Nn);
    c.conn.assign(m->conn, m->conn + 8 * (size_t)c.Ne);
    for (uint32_t v : c.conn)
        if (v >= c.Nn) {
            err = "connectivity references node " + std::to_string(v) + " >= n_nodes";
            return false;
        }
    for (double v : c.x0)
        if (!std::isfinite(v)) {
            err = "non-finite node coordinate";
            return false;
        }
    c.v_init.assign(3 * (size_t)c.Nn, 0.0);
    if (m->v_init) c.v_init.assign(m->v_init, m->v_init + 3 * (size_t)c.Nn);
    for (double v : c.v_init)
        if (!std::isfinite(v)) {
            err = "non-finite initial velocity";
            return false;
        }
    c.E = m->E;
    c.nu = m->nu;
    c.rho = m->rho;
    if (!(c.rho > 0) || !(c.E > 0) || !(c.nu > -1.0 && c.nu < 0.5)) {
        err = "invalid material (need rho > 0, E > 0, -1 < nu < 0.5)";
        return false;
    }
    c.plastic = m->plastic;
    c.yield0 = m->yield0;
    c.hard = m->hard;
    c.c_dr = m->c_dr;
    c.c_cr = m->c_cr;
    c.factor = m->factor;
    c.cvisl = m->cvisl;
    c.cvisq = m->cvisq;
    c.total_time = m->total_time;
    c.steps.resize(m->n_steps);
    for (uint32_t s = 0; s < m->n_steps; ++s) {
        const ensc_step_def& sd = m->steps[s];
        StepCopy& sc = c.steps[s];
        sc.period = sd.period;
        sc.reset_spc = sd.reset_spc;
        sc.reset_vel = sd.reset_vel;
        sc.reset_dload = sd.reset_dload;
        sc.grav_on = sd.grav_on;
        sc.grav_value = sd.grav_value;
        sc.grav_amp = sd.grav_amp;
        for (int d = 0; d < 3; ++d) sc.grav_dir[d] = sd.grav_dir[d];
        auto cp = [&](const ensc_bc_entry* b, uint32_t n, std::vector<BcCopy>& out) -> bool {
            for (uint32_t i = 0; i < n; ++i) {
                BcCopy bc;
                bc.nodes.assign(b[i].nodes, b[i].nodes + b[i].n_nodes);
                for (uint32_t v : bc.nodes)
                    if (v >= c.Nn) {
                        err = "BC references node " + std::to_string(v) + " >= n_nodes";
                        return false;
                    }
                if (b[i].d0 < 0 || b[i].d1 > 2) {
                    err = "BC dof range outside 0..2";
                    return false;
                }
                bc.d0 = b[i].d0;
                bc.d1 = b[i].d1;
                bc.cond = b[i].cond;
                bc.value = b[i].value;
                bc.amp = b[i].amp;
                out.push_back(std::move(bc));
            }
            return true;
        };
        if (!cp(sd.spc, sd.n_spc, sc.spc) || !cp(sd.vel, sd.n_vel, sc.vel)) return false;
        sc.reset_dsload = sd.reset_dsload;
        for (uint32_t i = 0; i < sd.n_dsload; ++i) {
            const ensc_dsload_entry& d = sd.dsload[i];
            DsloadCopy dc;
            dc.elems.assign(d.elems, d.elems + d.n_faces);
            dc.faces.assign(d.faces, d.faces + d.n_faces);
            for (size_t q = 0; q < dc.elems.size(); ++q)
                if (dc.elems[q] >= c.Ne || dc.faces[q] < 0 || dc.faces[q] > 5) {
                    err = "*Dsload face references element " + std::to_string(dc.elems[q]) + " face " +
                          std::to_string(dc.faces[q]) + " (need element < n_elems, face 0..5)";
                    return false;
                }
            dc.amp = d.amp;
            dc.value = d.value;
            sc.dsload.push_back(std::move(dc));
        }
    }
    c.amp_off.push_back(0);
    for (uint32_t i = 0; i < m->n_amps; ++i) {
        for (uint32_t q = 0; q < m->amps[i].n; ++q) {
            c.amp_t.push_back(m->amps[i].t[q]);
            c.amp_a.push_back(m->amps[i].a[q]);
        }
        c.amp_off.push_back((int32_t)c.amp_t.size());
    }
    // Material_elastic::update (Material_elastic.cpp:9-14)
    c.G = c.E / ((1.0 + c.nu) * 2.0);
    c.K = c.E / (1.0 - 2.0 * c.nu) / 3.0;
    c.lambda = 2.0 * c.G * c.nu / (1.0 - 2.0 * c.nu);
    c.wos = std::sqrt((c.lambda + 2.0 * c.G) / c.rho);
    return true;
}

template <class T>
static ensc_status upload(T** dst, const std::vector<T>& src, size_t n_alloc, DeviceArena& arena) {
    const size_t bytes = std::max<size_t>(n_alloc, 1) * sizeof(T);
    void* p = arena.alloc(bytes);
    if (!p) return fail(ENSC_ERR_MEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    *dst = static_cast<T*>(p);
    HIPCHK(hipMemset(p, 0, bytes));
    if (!src.empty()) HIPCHK(hipMemcpy(p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return ENSC_OK;
}

void* DeviceArena::alloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, round_up(bytes, 256)) != hipSuccess) return nullptr;
    ptrs.push_back(p);
    total += round_up(bytes, 256);
    return p;
}
DeviceArena::~DeviceArena() {
    for (void* p : ptrs) (void)hipFree(p);
}

// Per-job precision of the mixed engine.  A quasi-static run - many wave
// transits, so the state changes little per step - keeps the per-step FP32
// errors of the element arithmetic in its stress, and on small, lightly
// restrained meshes they grow into the displacement: on the reference's own
// decks (810-7500 transits) FP32 element arithmetic reaches 1e-4 to 3e-3,
// with the stress and QIA accumulated exactly or not (measured on a CPU
// emulation and on the GPU), while dynamic runs of a few hundred transits
// stay near 1e-7.  Such jobs run the FP64 kernels.  The rule reads the job
// alone: total time x wave speed over the largest extent of the mesh.
int job_precision(const ModelCopy& m, int engine_precision, int fp64_rule) {
    if (engine_precision != ENSC_MIXED || fp64_rule == 2) return ENSC_FP64;
    if (fp64_rule == 1) return ENSC_MIXED;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (size_t i = 0; i < m.Nn; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], m.x0[3 * i + d]);
            hi[d] = std::max(hi[d], m.x0[3 * i + d]);
        }
    const double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    const bool quasi_static = ext > 0.0 && m.total_time * m.wos / ext >= QUASI_STATIC_TRANSITS;
    return quasi_static ? ENSC_FP64 : ENSC_MIXED;
}

// Host half of building a job: setup arithmetic, numbering and every table
// the device needs, into J.img.  Touches nothing but J (no HIP calls), so
// it may run on any thread.
ensc_status prepare_job(Job& J, int engine_precision, int reorder, int blk_e, int blk_n, int fp64_rule) {
    const ModelCopy& M = J.model;
    const size_t Ne = M.Ne, Nn = M.Nn;
    const int precision = job_precision(M, engine_precision, fp64_rule);
    J.precision = precision;
    const bool mixed = precision == ENSC_MIXED;
    std::string err;
    std::vector<double> G8, V0, invM;
    if (!setup_geometry(M.x0, M.conn, Ne, M.rho, G8, V0, invM, Nn, err)) return fail(ENSC_ERR_ARG, err);
    // numbering
    if (reorder) {
        morton_order(M.x0, M.conn, Ne, Nn, J.e_new2old, J.n_new2old);
    } else {
        J.e_new2old.resize(Ne);
        J.n_new2old.resize(Nn);
        std::iota(J.e_new2old.begin(), J.e_new2old.end(), 0u);
        std::iota(J.n_new2old.begin(), J.n_new2old.end(), 0u);
    }
    J.n_old2new.assign(Nn, 0);
    for (size_t i = 0; i < Nn; ++i) J.n_old2new[J.n_new2old[i]] = (uint32_t)i;
    J.e_old2new.assign(Ne, 0);
    for (size_t i = 0; i < Ne; ++i) J.e_old2new[J.e_new2old[i]] = (uint32_t)i;
    const size_t es = round_up(Ne, 256), ns = round_up(Nn, 256);
    // Device code addresses every per-job array with a 32-bit byte offset
    // from the array's base, so each must stay under 4 GiB: the per-corner
    // forces (24 values per element) and the node records bound the size of
    // one job (FP64: ~22M elements, mixed: ~44M; larger meshes would be split
    // across engines or GPUs).
    {
        const size_t R = mixed ? sizeof(float) : sizeof(double);
        const size_t largest = std::max({24 * R * es, (size_t)VU_REC * sizeof(double) * ns,
                                         (size_t)POS_REC * sizeof(double) * ns});
        if (Ne >= (1u << 29) || largest >= ((size_t)1 << 32))
            return fail(ENSC_ERR_ARG, "mesh too large for one engine job: " + std::to_string(Ne) + " elements, " +
                                          std::to_string(Nn) + " nodes (a per-job array would reach 4 GiB)");
    }
    J.ne_s = es;
    J.nn_s = ns;
    // element SoA arrays in the new numbering
    std::vector<uint32_t> conn_d(8 * es, 0);
    for (size_t i = 0; i < Ne; ++i) {
        const size_t eo = J.e_new2old[i];
        for (int k = 0; k < 8; ++k) conn_d[(size_t)k * es + i] = J.n_old2new[M.conn[8 * eo + k]];
    }
    // ELL gather table: per node the (new element, corner) slots sorted by
    // (corner, original element id) - the reference's summation order.
    std::vector<uint32_t> cnt(Nn, 0);
    for (size_t e = 0; e < Ne; ++e)
        for (int k = 0; k < 8; ++k) cnt[M.conn[8 * e + k]]++;
    uint32_t W = 0;
    for (uint32_t v : cnt) W = std::max(W, v);
    if ((size_t)W * ns * sizeof(uint32_t) >= ((size_t)1 << 32))
        return fail(ENSC_ERR_ARG, "a node is shared by " + std::to_string(W) +
                                      " elements: the gather table would reach 4 GiB");
    J.ell_w = W;
    std::vector<uint32_t> ell((size_t)W * ns, ELL_EMPTY);
    std::vector<uint32_t> fillc(Nn, 0);
    for (int k = 0; k < 8; ++k)
        for (size_t eo = 0; eo < Ne; ++eo) {
            const uint32_t no = M.conn[8 * eo + k];
            const uint32_t nn_ = J.n_old2new[no];
            ell[(size_t)fillc[no]++ * ns + nn_] = (J.e_old2new[eo] << 3) | (uint32_t)k;
        }
    // LDS staging lists: per element block, its distinct nodes (ascending new
    // ids) and each corner's 16-bit index into that list.
    std::vector<uint32_t> blk_ptr, blk_nodes;
    std::vector<uint16_t> lconn(8 * es, 0);
    J.lds_nodes = 0;
    {
        const size_t nb = (Ne + blk_e - 1) / blk_e;
        blk_ptr.assign(nb + 1, 0);
        std::vector<uint32_t> tmp;
        for (size_t b = 0; b < nb; ++b) {
            const size_t e0 = b * blk_e, e1 = std::min(Ne, e0 + blk_e);
            tmp.clear();
            for (size_t e = e0; e < e1; ++e)
                for (int k = 0; k < 8; ++k) tmp.push_back(conn_d[(size_t)k * es + e]);
            std::sort(tmp.begin(), tmp.end());
            tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
            for (size_t e = e0; e < e1; ++e)
                for (int k = 0; k < 8; ++k) {
                    const uint32_t n = conn_d[(size_t)k * es + e];
                    lconn[(size_t)k * es + e] = (uint16_t)(std::lower_bound(tmp.begin(), tmp.end(), n) - tmp.begin());
                }
            blk_nodes.insert(blk_nodes.end(), tmp.begin(), tmp.end());
            blk_ptr[b + 1] = (uint32_t)blk_nodes.size();
            J.lds_nodes = std::max<uint32_t>(J.lds_nodes, (uint32_t)tmp.size());
        }
    }
    // Block assembly needs every element block to be free of corner
    // conflicts: no two of its elements share a node at the same corner
    // index.  Hex meshes of one orientation always are.
    J.asm_capable = true;
    {
        const size_t nb = blk_ptr.size() - 1;
        std::vector<uint32_t> stamp(J.lds_nodes + 1, 0);
        uint32_t tag = 0;
        for (size_t b = 0; b < nb && J.asm_capable; ++b) {
            const size_t e0 = b * blk_e, e1 = std::min(Ne, e0 + blk_e);
            for (int k = 0; k < 8 && J.asm_capable; ++k) {
                ++tag;
                for (size_t e = e0; e < e1; ++e) {
                    const uint16_t l = lconn[(size_t)k * es + e];
                    if (stamp[l] == tag) {
                        J.asm_capable = false;
                        break;
                    }
                    stamp[l] = tag;
                }
            }
        }
    }
    // Block assembly adds corner k's forces in phase k = 0..7; the first
    // (element, corner) of a block to reach a node in that order stores its
    // force instead of adding it, so the accumulators need no clearing.  It
    // is marked by bit 15 of the corner's slot (slots stay below 2^15).
    if (J.asm_capable) {
        const size_t nb = blk_ptr.size() - 1;
        std::vector<uint8_t> seen(J.lds_nodes + 1, 0);
        for (size_t b = 0; b < nb; ++b) {
            const size_t e0 = b * blk_e, e1 = std::min(Ne, e0 + blk_e);
            std::fill(seen.begin(), seen.end(), 0);
            for (int k = 0; k < 8; ++k)
                for (size_t e = e0; e < e1; ++e) {
                    uint16_t& l = lconn[(size_t)k * es + e];
                    if (!seen[l]) {
                        seen[l] = 1;
                        l |= 0x8000u;
                    }
                }
        }
    }
    // each node's positions in blk_nodes (one per element block holding it),
    // ascending: the order the nodal update adds the block partials
    std::vector<uint32_t> pos_ptr(Nn + 1, 0), pos_list(blk_nodes.size());
    for (uint32_t g : blk_nodes) pos_ptr[g + 1]++;
    for (size_t i = 0; i < Nn; ++i) pos_ptr[i + 1] += pos_ptr[i];
    {
        std::vector<uint32_t> fill(pos_ptr.begin(), pos_ptr.end() - 1);
        for (size_t p = 0; p < blk_nodes.size(); ++p) pos_list[fill[blk_nodes[p]]++] = (uint32_t)p;
    }
    // node records
    std::vector<double> pos((size_t)Nn * POS_REC, 0.0), vu((size_t)Nn * VU_REC, 0.0);
    for (size_t i = 0; i < Nn; ++i) {
        const size_t o = J.n_new2old[i];
        for (int d = 0; d < 3; ++d) {
            pos[i * POS_REC + d] = M.x0[3 * o + d];
            vu[i * VU_REC + d] = M.v_init[3 * o + d];
        }
        pos[i * POS_REC + 3] = invM[o];
    }
    // initial kinetic energy
    double ke0 = 0.0;
    for (size_t n = 0; n < Nn; ++n) {
        double k2 = 0.0;
        for (int d = 0; d < 3; ++d) k2 += M.v_init[3 * n + d] * M.v_init[3 * n + d];
        ke0 += 0.5 * k2 / invM[n];
    }
    J.ke0 = ke0;
    JobImage& I = *(J.img = std::unique_ptr<JobImage>(new JobImage()));
    I.precision = precision;
    I.blk_e = blk_e;
    I.blk_n = blk_n;
    I.W = W;
    I.es = es;
    I.ns = ns;
    // the mixed engine rebuilds Gi and V0 from the reference positions in
    // its element body; only the FP64 engine stores them
    if (!mixed) {
        I.gi64.assign(9 * es, 0.0);
        I.v064.assign(es, 0.0);
        for (size_t i = 0; i < Ne; ++i) {
            const size_t eo = J.e_new2old[i];
            for (int q = 0; q < 9; ++q) I.gi64[(size_t)q * es + i] = G8[9 * eo + q];
            I.v064[i] = V0[eo];
        }
    }
    I.conn_d = std::move(conn_d);
    I.ell = std::move(ell);
    // device layout: corners 2p and 2p + 1 of an element side by side, so
    // one 32-bit load reads a pair ([4][ne_s] words)
    I.lconn.assign(8 * es, 0);
    for (size_t k = 0; k < 8; ++k)
        for (size_t e = 0; e < es; ++e) I.lconn[((k >> 1) * es + e) * 2 + (k & 1)] = lconn[k * es + e];
    I.blk_ptr = std::move(blk_ptr);
    I.blk_nodes = std::move(blk_nodes);
    I.pos_ptr = std::move(pos_ptr);
    I.pos_list = std::move(pos_list);
    I.pos = std::move(pos);
    I.vu = std::move(vu);
    J.blk_ptr_h = I.blk_ptr;
    return ENSC_OK;
}

// Device half: allocations and uploads from J.img, then the image is freed.
// The job's ctl/partials pointers are filled in by the engine (they live in
// engine-wide arrays).
ensc_status upload_job(Job& J) {
    if (!J.img) return fail(ENSC_ERR_STATE, "job has no prepared image");
    const JobImage& I = *J.img;
    const ModelCopy& M = J.model;
    const size_t Ne = M.Ne, Nn = M.Nn, es = I.es, ns = I.ns;
    const bool mixed = I.precision == ENSC_MIXED;
    const int blk_e = I.blk_e, blk_n = I.blk_n;
    const uint32_t W = I.W;
    DeviceArena& A = J.arena;
    JobDev& D = J.dev;
    std::memset(&D, 0, sizeof(D));
    ensc_status st;
    uint32_t* d_conn;
    if ((st = upload(&d_conn, I.conn_d, I.conn_d.size(), A))) return st;
    D.conn = d_conn;
    if (mixed) {
        float *ds, *dq, *dp, *df;
        std::vector<float> none;
        if ((st = upload(&ds, none, 6 * es, A))) return st;
        if ((st = upload(&dq, none, 12 * es, A))) return st;
        if ((st = upload(&dp, none, es, A))) return st;
        if ((st = upload(&df, none, 8 * es * 3, A))) return st;
        uint32_t* dc;
        std::vector<uint32_t> no_c;
        if ((st = upload(&dc, no_c, Nn, A))) return st;
        D.ucomp = dc;
        float* dpf;
        if ((st = upload(&dpf, none, I.blk_nodes.size() * 3, A))) return st;
        D.part_f = dpf;
        D.Gi = nullptr;
        D.V0 = nullptr;
        D.sigma = ds;
        D.qia = dq;
        D.epsp = dp;
        D.fel = df;
        std::vector<double> im(Nn);
        for (size_t i = 0; i < Nn; ++i) im[i] = I.pos[i * POS_REC + 3];
        double* dim;
        if ((st = upload(&dim, im, im.size(), A))) return st;
        D.inv_m = dim;
    } else {
        double *db, *dv, *ds, *dq, *dp, *df;
        if ((st = upload(&db, I.gi64, I.gi64.size(), A))) return st;
        if ((st = upload(&dv, I.v064, I.v064.size(), A))) return st;
        std::vector<double> none;
        if ((st = upload(&ds, none, 6 * es, A))) return st;
        if ((st = upload(&dq, none, 12 * es, A))) return st;
        if ((st = upload(&dp, none, es, A))) return st;
        if ((st = upload(&df, none, 8 * es * 3, A))) return st;
        D.Gi = db;
        D.V0 = dv;
        D.sigma = ds;
        D.qia = dq;
        D.epsp = dp;
        D.fel = df;
        double* dpf;
        if ((st = upload(&dpf, none, I.blk_nodes.size() * 3, A))) return st;
        D.part_f = dpf;
    }
    // block assembly: positions of each node's partials
    {
        uint32_t *dpp, *dpl;
        if ((st = upload(&dpp, I.pos_ptr, I.pos_ptr.size(), A))) return st;
        if ((st = upload(&dpl, I.pos_list, I.pos_list.size(), A))) return st;
        D.pos_ptr = dpp;
        D.pos_list = dpl;
    }
    double *d_pos, *d_vu;
    if ((st = upload(&d_pos, I.pos, I.pos.size(), A))) return st;
    if ((st = upload(&d_vu, I.vu, I.vu.size(), A))) return st;
    D.pos = d_pos;
    D.vu = d_vu;
    std::vector<double> none_d;
    if ((st = upload(&D.rf, none_d, 3 * Nn, A))) return st;
    std::vector<uint8_t> none_u8;
    uint8_t* d_mask;
    if ((st = upload(&d_mask, none_u8, Nn, A))) return st;
    D.bcmask = d_mask;
    J.d_bcmask = d_mask;
    double* d_bcval;
    if ((st = upload(&d_bcval, none_d, 3 * Nn, A))) return st;
    D.bcval = d_bcval;
    J.d_bcval = d_bcval;
    std::vector<int16_t> none_i16;
    int16_t* d_bcamp;
    if ((st = upload(&d_bcamp, none_i16, 3 * Nn, A))) return st;
    D.bcamp = d_bcamp;
    J.d_bcamp = d_bcamp;
    uint32_t* d_ell;
    if ((st = upload(&d_ell, I.ell, I.ell.size(), A))) return st;
    D.ell = d_ell;
    uint16_t* d_lconn;
    uint32_t *d_bptr, *d_bnodes;
    if ((st = upload(&d_lconn, I.lconn, I.lconn.size(), A))) return st;
    if ((st = upload(&d_bptr, I.blk_ptr, I.blk_ptr.size(), A))) return st;
    if ((st = upload(&d_bnodes, I.blk_nodes, I.blk_nodes.size(), A))) return st;
    D.lconn = d_lconn;
    D.blk_ptr = d_bptr;
    D.blk_nodes = d_bnodes;
    double *d_at, *d_aa;
    int32_t* d_ao;
    if ((st = upload(&d_at, M.amp_t, M.amp_t.size(), A))) return st;
    if ((st = upload(&d_aa, M.amp_a, M.amp_a.size(), A))) return st;
    if ((st = upload(&d_ao, M.amp_off, M.amp_off.size(), A))) return st;
    D.amp_t = d_at;
    D.amp_a = d_aa;
    D.amp_off = d_ao;
    D.n_amp = (int32_t)M.amp_off.size() - 1;
    D.lambda = M.lambda;
    D.G = M.G;
    D.K = M.K;
    D.rho = M.rho;
    D.wos = M.wos;
    D.c_dr = M.c_dr;
    D.c_cr = M.c_cr;
    D.cvisl = M.cvisl;
    D.cvisq = M.cvisq;
    D.factor = M.factor;
    D.yield0 = M.yield0;
    D.hard = M.hard;
    D.mc.lam = (float)M.lambda;
    D.mc.two_g = (float)(2.0 * M.G);
    D.mc.three_g = (float)(3.0 * M.G);
    D.mc.inv_ret = (float)(1.0 / (3.0 * M.G + M.hard));
    D.mc.yield0 = (float)M.yield0;
    D.mc.hard = (float)M.hard;
    D.mc.dr = (float)(M.c_dr * (4.0 / 3.0 * M.G + M.K) * 0.1 / std::sqrt(8.0));
    D.mc.cr = (float)(M.c_cr * M.rho * M.wos / 4.0);
    D.mc.rho = (float)M.rho;
    D.mc.cq2 = (float)(M.cvisq * M.cvisq);
    D.mc.clw = (float)(M.cvisl * M.wos);
    D.mc.pad = 0.0f;
    D.mc.inv_wos = 1.0 / M.wos;
    D.mc.cq2_wos = M.cvisq * M.cvisq / M.wos;
    D.plastic = M.plastic;
    D.ell_w = (int32_t)W;
    D.ne = (uint32_t)Ne;
    D.nn = (uint32_t)Nn;
    D.ne_s = (uint32_t)es;
    D.nn_s = (uint32_t)ns;
    D.nblk_e = (uint32_t)((Ne + blk_e - 1) / blk_e);
    D.nblk_n = (uint32_t)((Nn + blk_n - 1) / blk_n);
    D.grav_amp = -1;
    // dyn_bc: any velocity BC anywhere with an amplitude
    for (const auto& s : M.steps)
        for (const auto& b : s.vel)
            if (b.amp >= 0) D.dyn_bc = 1;
    // element blocks holding a node that some step constrains (velocity or
    // SPC lines): the blocks whose staged velocities the pipelined pass
    // overrides on such jobs' step ticks
    if (D.dyn_bc) {
        std::vector<uint8_t> blk_bc(D.nblk_e, 0);
        std::vector<uint8_t> bc(Nn, 0);
        for (const auto& s : M.steps) {
            for (const auto& b : s.vel)
                for (uint32_t no : b.nodes) bc[J.n_old2new[no]] = 1;
            for (const auto& b : s.spc)
                if (b.cond == 1)
                    for (uint32_t no : b.nodes) bc[J.n_old2new[no]] = 1;
        }
        for (uint32_t k = 0; k < D.nblk_e; ++k)
            for (uint32_t q = I.blk_ptr[k]; q < I.blk_ptr[k + 1]; ++q)
                if (bc[I.blk_nodes[q]]) {
                    blk_bc[k] = 1;
                    break;
                }
        uint8_t* d_bbc;
        if ((st = upload(&d_bbc, blk_bc, blk_bc.size(), A))) return st;
        D.blk_bc = d_bbc;
    }
    J.grav.resize(M.steps.size());
    for (size_t s = 0; s < M.steps.size(); ++s) {
        J.grav[s].on = M.steps[s].grav_on;
        J.grav[s].amp = M.steps[s].grav_amp;
        J.grav[s].g = M.steps[s].grav_value;
        for (int d = 0; d < 3; ++d) J.grav[s].dir[d] = M.steps[s].grav_dir[d];
    }
    J.img.reset();
    return ENSC_OK;
}

// setCurrentStep's BC inheritance (exDyna3D.cpp:1760-1908), on the job's
// host copy; returns false when the reference would return early.
bool job_set_step(Job& J, uint32_t si) {
    const ModelCopy& M = J.model;
    if (si >= M.steps.size()) return false;
    if (si == J.cur_step) return false;
    if (si > 0) {
        J.prev_spc = J.cur_spc;
        J.prev_vel = J.cur_vel;
    }
    J.cur_step = si;
    const StepCopy& st = M.steps[si];
    auto ids = [](uint32_t s, const std::vector<BcCopy>& v, std::vector<BcRef>& out) {
        for (uint32_t i = 0; i < v.size(); ++i) out.push_back({s, i});
    };
    J.cur_spc.clear();
    J.cur_vel.clear();
    if (si == 0) {
        ids(0, st.spc, J.cur_spc);
        ids(0, st.vel, J.cur_vel);
    } else {
        ids(si, st.spc, J.cur_spc);
        if (!st.reset_spc) {
            if (!J.prev_spc.empty())
                J.cur_spc.insert(J.cur_spc.end(), J.prev_spc.begin(), J.prev_spc.end());
            else
                ids(si - 1, M.steps[si - 1].spc, J.cur_spc);
        }
        ids(si, st.vel, J.cur_vel);
        if (!st.reset_vel) J.cur_vel.insert(J.cur_vel.end(), J.prev_vel.begin(), J.prev_vel.end());
        if (!st.reset_dload && !J.grav[si].on && J.grav[si - 1].on) J.grav[si] = J.grav[si - 1];
        // *Dsload lines (exDyna3D.cpp:1897-1907): the step's own lines; with
        // none, the previous step's unless op=NEW
        if (!st.dsload.empty())
            J.dsl_step = (int32_t)si;
        else if (st.reset_dsload)
            J.dsl_step = -1;
    }
    return true;
}

// C3D8 face k (S(k+1)) as element corners, ordered so that the right-hand
// rule gives the normal pointing into the element (the face numbering of
// Element_HexN8.cpp:50-89's node order: S1 = 1-2-3-4, S2 = 5-8-7-6,
// S3 = 1-5-6-2, S4 = 2-6-7-3, S5 = 3-7-8-4, S6 = 4-8-5-1).
static const int FACE_CORNERS[6][4] = {{0, 1, 2, 3}, {4, 7, 6, 5}, {0, 4, 5, 1},
                                       {1, 5, 6, 2}, {2, 6, 7, 3}, {3, 7, 4, 0}};

// The loaded faces of the job's current *Dsload lines (new numbering): per
// face its 4 nodes, pressure and amplitude, and per node its faces in
// ascending face order (the order the nodal update adds their forces).
void job_dsload_arrays(const Job& J, std::vector<uint32_t>& fnodes, std::vector<double>& fval,
                       std::vector<int32_t>& famp, std::vector<uint32_t>& nptr, std::vector<uint32_t>& nlist) {
    const ModelCopy& M = J.model;
    fnodes.clear();
    fval.clear();
    famp.clear();
    if (J.dsl_step >= 0)
        for (const DsloadCopy& d : M.steps[J.dsl_step].dsload)
            for (size_t q = 0; q < d.elems.size(); ++q) {
                for (int c = 0; c < 4; ++c)
                    fnodes.push_back(J.n_old2new[M.conn[8 * (size_t)d.elems[q] + FACE_CORNERS[d.faces[q]][c]]]);
                fval.push_back(d.value);
                famp.push_back(d.amp);
            }
    nptr.assign((size_t)M.Nn + 1, 0);
    for (uint32_t n : fnodes) nptr[n + 1]++;
    for (size_t i = 0; i < M.Nn; ++i) nptr[i + 1] += nptr[i];
    nlist.assign(fnodes.size(), 0);
    std::vector<uint32_t> fill(nptr.begin(), nptr.end() - 1);
    for (size_t i = 0; i < fnodes.size(); ++i) nlist[fill[fnodes[i]]++] = (uint32_t)(i / 4);
}

// Per-node BC arrays for the job's current BC lists (new numbering).
void job_bc_arrays(const Job& J, std::vector<uint8_t>& mask, std::vector<double>& val, std::vector<int16_t>& amp,
                   std::vector<uint32_t>& nodes) {
    const ModelCopy& M = J.model;
    mask.assign(M.Nn, 0);
    val.assign(3 * (size_t)M.Nn, 0.0);
    amp.assign(3 * (size_t)M.Nn, (int16_t)-1);
    for (const BcRef& r : J.cur_vel) {
        const BcCopy& b = M.steps[r.step].vel[r.idx];
        for (uint32_t no : b.nodes) {
            const uint32_t n = J.n_old2new[no];
            for (int d = b.d0; d <= b.d1; ++d) {
                mask[n] |= (uint8_t)(BC_VEL0 << d);
                val[3 * (size_t)n + d] = b.value;  // later lines overwrite, as sequential application would
                amp[3 * (size_t)n + d] = (int16_t)b.amp;
            }
        }
    }
    for (const BcRef& r : J.cur_spc) {
        const BcCopy& b = M.steps[r.step].spc[r.idx];
        if (b.cond != 1) continue;
        for (uint32_t no : b.nodes) {
            const uint32_t n = J.n_old2new[no];
            for (int d = b.d0; d <= b.d1; ++d) mask[n] |= (uint8_t)(BC_SPC0 << d);
        }
    }
    nodes.clear();
    for (uint32_t n = 0; n < M.Nn; ++n)
        if (mask[n]) nodes.push_back(n);
}

// ---- fluid-structure coupling: the virtual particles ----
//
// exDyna3D::get_fsi_sph_Ele_face_index (exDyna3D.cpp:1282-1361): a node is
// on the FSI boundary when fewer than 8 elements hold it; an element face
// is an FSI face when its 4 corners all are.  exDyna3D::processInteraction-
// Elements and generateVirtualParticlesForFace (1428-1638): per element in
// ascending order, per FSI face in face order, n_layers particles step
// inward from the face centre along the face's unit normal, `dist` apart
// in space: the unit step is dist / |J n|, J = dN(face centre) X the 3x3
// product of the shape-function derivatives with the node positions and n
// the unit normal (the reference's product J n, not J^T n).  The face
// numbering is Element_HexN8::Ele_face_indices (Element_HexN8.cpp:21-26):
// faces 0..5 are unit-coordinate faces xi = -1, xi = +1, eta = -1, eta = +1,
// zeta = -1, zeta = +1.
static const int FSI_FACE_CORNERS[6][4] = {{3, 0, 4, 7}, {1, 2, 6, 5}, {0, 1, 5, 4},
                                           {2, 3, 7, 6}, {1, 0, 3, 2}, {4, 5, 6, 7}};

// the corner's unit coordinates (Element_HexN8.cpp:46-93)
static void hex_corner(int k, double c[3]) {
    c[0] = ((k + 1) & 2) ? 1.0 : -1.0;
    c[1] = (k & 2) ? 1.0 : -1.0;
    c[2] = (k & 4) ? 1.0 : -1.0;
}

void fsi_particles(const ModelCopy& M, const std::vector<double>& x, int n_layers, double dist, FsiParticles& out) {
    out.elems.clear();
    out.eptr.assign(1, 0);
    out.unit.clear();
    std::vector<uint32_t> held(M.Nn, 0);
    for (size_t q = 0; q < 8 * (size_t)M.Ne; ++q) held[M.conn[q]]++;
    auto on_boundary = [&](uint32_t n) { return held[n] > 0 && held[n] < 8; };
    for (uint32_t e = 0; e < M.Ne; ++e) {
        const uint32_t* cn = &M.conn[8 * (size_t)e];
        const size_t before = out.unit.size();
        for (int f = 0; f < 6; ++f) {
            bool all = true;
            for (int c = 0; c < 4; ++c) all = all && on_boundary(cn[FSI_FACE_CORNERS[f][c]]);
            if (!all) continue;
            // face centre and inward unit normal in unit coordinates
            double pc[3] = {0.0, 0.0, 0.0}, nu[3] = {0.0, 0.0, 0.0};
            pc[f / 2] = (f & 1) ? 1.0 : -1.0;
            nu[f / 2] = (f & 1) ? -1.0 : 1.0;
            // J(i, j) = sum_k dN_k/dxi_i (face centre) X_k(j)
            double J[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double s = 0.0;
                    for (int k = 0; k < 8; ++k) {
                        double c[3];
                        hex_corner(k, c);
                        double dn;
                        if (i == 0)
                            dn = 1.0 / 8.0 * c[0] * (1 + c[1] * pc[1]) * (1 + c[2] * pc[2]);
                        else if (i == 1)
                            dn = 1.0 / 8.0 * (1 + c[0] * pc[0]) * c[1] * (1 + c[2] * pc[2]);
                        else
                            dn = 1.0 / 8.0 * (1 + c[0] * pc[0]) * (1 + c[1] * pc[1]) * c[2];
                        const double t = dn * x[3 * (size_t)cn[k] + j];
                        s = k ? s + t : t;
                    }
                    J[i][j] = s;
                }
            double s2 = 0.0;
            for (int i = 0; i < 3; ++i) {
                const double ni = J[i][0] * nu[0] + J[i][1] * nu[1] + J[i][2] * nu[2];
                s2 = i ? s2 + ni * ni : ni * ni;
            }
            if (!(s2 > 1e-24)) continue;  // degenerate face: no particles (1475-1481)
            const double step = dist / std::sqrt(s2);
            for (int l = 0; l < n_layers; ++l)
                for (int d = 0; d < 3; ++d) out.unit.push_back(pc[d] + l * step * nu[d]);
        }
        if (out.unit.size() > before) {
            out.elems.push_back(e);
            out.eptr.push_back((uint32_t)(out.unit.size() / 3));
        }
    }
}

}  // namespace ensc
