// CPU FP64 restatement of the reference's central-difference step loop.
//
// TEST INFRASTRUCTURE: this file is the oracle the HIP engine is checked
// against, and the CPU baseline bench.py times.  The engine never calls it.
//
// It restates ssadz/EnSCDynamics Explicit3D/source/exDyna3D.cpp in the same
// arithmetic order (each function cites the lines it follows).  Where the
// reference races (its 8 colour arrays when two elements share a node at the
// same corner, exDyna3D.cpp:584-634; overlapping BC tasks, 381-489) this
// restatement fixes one order: per node, contributions are grouped by corner
// 0..7 and summed within a group in ascending element id, which is the
// reference's order whenever each (node, corner) pair is unique (structured
// meshes).  Eigen's 3x3 determinant and inverse are restated from
// eigen/include/eigen3/Eigen/src/LU/Determinant.h:19-24,56-65 and
// InverseImpl.h:127-176.  Small matrix products sum their inner index in
// ascending order.
//
// Extensions with no reference function (pinned by closed forms in tests/):
//  * J2 radial return with linear isotropic hardening after the Jaumann
//    update (Simo & Hughes, Computational Inelasticity, 1998, Box 3.1).
//  * Energy sums (internal, hourglass, bulk-viscous, kinetic, external).
//    External work is the reactions' work at velocity-BC dofs, gravity's
//    work on every dof and the kinetic energy constraints add or remove
//    when they set v, so KE + IE + HG + VIS - W_ext = KE0 closes.
//  * An amplitude on velocity BCs (piecewise-linear in total time).
//  * Surface pressures (*Dsload P lines, which the reference parses and
//    inherits across steps, exDyna3D.cpp:1897-1907, but never applies): each
//    loaded face puts p amp A / 4 on each of its 4 nodes, A the face's vector
//    area (into the element) on the positions at the start of the step, amp
//    averaged over the step like gravity's; the forces are added after
//    gravity, per node in ascending face order, and their work enters the
//    external work like gravity's.
//  * The external work of the virtual particles' forces (the FSI coupling
//    itself, exDyna3D.cpp:1281-1758, is restated below).
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace {

struct Gravity {
    bool on = false;
    int amp = -1;
    double value = 0, dir[3] = {0, 0, 0};
};

struct Dsload {
    std::vector<uint32_t> elems;
    std::vector<int32_t> faces;
    int amp = -1;
    double value = 0;
};

// C3D8 face k (S(k+1)) as element corners, ordered so that the right-hand
// rule gives the normal into the element: S1 = 1-2-3-4, S2 = 5-8-7-6,
// S3 = 1-5-6-2, S4 = 2-6-7-3, S5 = 3-7-8-4, S6 = 4-8-5-1 in the corner
// numbering of Element_HexN8.cpp:50-89.
const int FACE_CORNERS[6][4] = {{0, 1, 2, 3}, {4, 7, 6, 5}, {0, 4, 5, 1}, {1, 5, 6, 2}, {2, 6, 7, 3}, {3, 7, 4, 0}};

struct Sim {
    // --- model ---
    size_t Nn = 0, Ne = 0;
    std::vector<uint32_t> conn;  // [Ne][8]
    double E = 0, nu = 0, rho = 0, G = 0, K = 0, lambda = 0, WOS = 0;
    bool plastic = false;
    double yield0 = 0, hard = 0;
    double c_dr = 0.013, c_cr = 0.05, factor = 0.71, Cvisl = 0.06, Cvisq = 1.2;
    const double volRateThreshold = 1e-10;  // exDyna3D.h:188
    std::vector<double> v_init;
    std::vector<ensc_step_def> steps;
    std::vector<Gravity> grav;  // per step, mutated by inheritance (exDyna3D.cpp:1889-1892)
    std::vector<std::vector<Dsload>> dsl;  // per step: its own *Dsload lines
    std::vector<int32_t> dsl_reset;
    std::vector<std::vector<double>> amp_t, amp_a;
    // --- state (exDyna3D.h:141-206) ---
    std::vector<double> x;             // [Nn][3] vertices
    std::vector<double> u, v, a, rf;   // [3Nn]
    std::vector<double> invM;          // [Nn] (the reference stores 3 equal copies)
    std::vector<double> BT;            // [Ne][8][3]
    std::vector<double> V0;            // [Ne]
    std::vector<double> sigma;         // [Ne][6]  xx yy zz xy xz yz
    std::vector<double> qia;           // [Ne][12]
    std::vector<double> epsp;          // [Ne]
    std::vector<double> dtStable, Length;
    std::vector<double> fel;           // [Ne][8][3] per-element nodal forces
    std::vector<double> Fel;           // [Ne][3][3] F of the last element pass (all_F, exDyna3D.h:153)
    std::vector<uint32_t> csr_ptr, csr_slot;  // node -> (8*e + k), sorted by (k, e)
    double time = 0, totalTime = 0, dt_i = 0, dt_i_1 = 0;
    size_t currentStepIndex = 0;
    std::vector<const ensc_bc_entry*> cur_spc, cur_vel, prev_spc, prev_vel;
    // --- energies (extension) ---
    double ke = 0, ie = 0, hg = 0, vis = 0, wext = 0, ke0 = 0;
    int64_t nsteps = 0;
    int nthreads = 1;
    std::vector<double> thr_part;  // per-thread partial sums

    double amp_interp(int id, double t) const {
        if (id < 0 || id >= (int)amp_t.size() || amp_t[id].empty()) return 1.0;
        const auto& T = amp_t[id];
        const auto& A = amp_a[id];
        if (t <= T.front()) return A.front();
        if (t >= T.back()) return A.back();
        size_t i = 1;
        while (i < T.size() && T[i] < t) ++i;
        const double w = (t - T[i - 1]) / (T[i] - T[i - 1]);
        return A[i - 1] + w * (A[i] - A[i - 1]);
    }

    // exDyna3D::deal_amp (exDyna3D.cpp:1222-1278): trapezoid average of the
    // amplitude over [time, time + dt_i].  Tables with fewer than 2 points
    // give 1 (the reference would index past an empty table).
    double deal_amp(int id) const {
        if (id < 0 || id >= (int)amp_t.size()) return 1.0;
        const auto& T = amp_t[id];
        const auto& A = amp_a[id];
        const size_t n = T.size();
        if (n < 2) return 1.0;
        const double dt = dt_i;
        double y = 1.0;
        for (size_t i = 0; i + 1 < n; ++i) {
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
            for (size_t j = 2; j + i < n; ++j) {
                const double xj = T[i + j], yj = A[i + j], xj_1 = T[i + j - 1], yj_1 = A[i + j - 1];
                if ((xm + dt / 2.0) > xj) {
                    A2 += (xj - xj_1) * (yj + yj_1) / 2.0;
                } else {
                    const double M = (yj - yj_1) / (xj - xj_1);
                    const double C = yj_1 - m * xj_1;  // the reference uses m here (1260)
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

    void load(const ensc_model_desc* m) {
        Nn = m->n_nodes;
        Ne = m->n_elems;
        x.assign(m->x0, m->x0 + 3 * Nn);
        conn.assign(m->conn, m->conn + 8 * Ne);
        E = m->E;
        nu = m->nu;
        rho = m->rho;
        plastic = m->plastic != 0;
        yield0 = m->yield0;
        hard = m->hard;
        c_dr = m->c_dr;
        c_cr = m->c_cr;
        factor = m->factor;
        Cvisl = m->cvisl;
        Cvisq = m->cvisq;
        totalTime = m->total_time;
        v_init.assign(m->v_init, m->v_init + 3 * Nn);
        steps.assign(m->steps, m->steps + m->n_steps);
        grav.resize(steps.size());
        for (size_t s = 0; s < steps.size(); ++s) {
            grav[s].on = steps[s].grav_on != 0;
            grav[s].amp = steps[s].grav_amp;
            grav[s].value = steps[s].grav_value;
            for (int d = 0; d < 3; ++d) grav[s].dir[d] = steps[s].grav_dir[d];
        }
        dsl.assign(steps.size(), {});
        dsl_reset.assign(steps.size(), 0);
        for (size_t s = 0; s < steps.size(); ++s) {
            dsl_reset[s] = steps[s].reset_dsload;
            for (uint32_t i = 0; i < steps[s].n_dsload; ++i) {
                const ensc_dsload_entry& d = steps[s].dsload[i];
                Dsload L;
                L.elems.assign(d.elems, d.elems + d.n_faces);
                L.faces.assign(d.faces, d.faces + d.n_faces);
                L.amp = d.amp;
                L.value = d.value;
                dsl[s].push_back(std::move(L));
            }
        }
        amp_t.resize(m->n_amps);
        amp_a.resize(m->n_amps);
        for (uint32_t i = 0; i < m->n_amps; ++i) {
            amp_t[i].assign(m->amps[i].t, m->amps[i].t + m->amps[i].n);
            amp_a[i].assign(m->amps[i].a, m->amps[i].a + m->amps[i].n);
        }
        // Material_elastic::update (Material_elastic.cpp:9-14)
        G = E / ((1.0 + nu) * 2.0);
        K = E / (1.0 - 2.0 * nu) / 3.0;
        lambda = 2.0 * G * nu / (1.0 - 2.0 * nu);
        WOS = std::sqrt((lambda + 2.0 * G) / rho);
        // CSR node -> (element, corner), canonical order (corner, element).
        csr_ptr.assign(Nn + 1, 0);
        for (size_t e = 0; e < Ne; ++e)
            for (int k = 0; k < 8; ++k) csr_ptr[conn[8 * e + k] + 1]++;
        for (size_t i = 0; i < Nn; ++i) csr_ptr[i + 1] += csr_ptr[i];
        csr_slot.assign(8 * Ne, 0);
        std::vector<uint32_t> fill(csr_ptr.begin(), csr_ptr.end() - 1);
        for (int k = 0; k < 8; ++k)
            for (size_t e = 0; e < Ne; ++e) {
                const uint32_t n = conn[8 * e + k];
                csr_slot[fill[n]++] = (uint32_t)(8 * e + k);
            }
        thr_part.assign((size_t)nthreads * 8, 0.0);
    }

    // exDyna3D::init_data + compute_invMassMatrix (exDyna3D.cpp:144-205,
    // 492-515; fevalues.h:233-265 at the single Gauss point xi=0, weight 8).
    void setup() {
        u.assign(3 * Nn, 0.0);
        v.assign(3 * Nn, 0.0);
        a.assign(3 * Nn, 0.0);
        rf.assign(3 * Nn, 0.0);
        sigma.assign(6 * Ne, 0.0);
        qia.assign(12 * Ne, 0.0);
        epsp.assign(Ne, 0.0);
        dtStable.assign(Ne, 0.0);
        Length.assign(Ne, 0.0);
        fel.assign(24 * Ne, 0.0);
        Fel.assign(9 * Ne, 0.0);
        BT.assign(24 * Ne, 0.0);
        V0.assign(Ne, 0.0);
        std::vector<double> mass(Nn, 0.0);
        // dN/dxi at xi = 0 for the corners of Element_HexN8.cpp:50-89.
        static const double cs[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                        {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
        double dN[3][8];
        for (int a8 = 0; a8 < 8; ++a8)
            for (int d = 0; d < 3; ++d) dN[d][a8] = 1.0 / 8.0 * cs[a8][d];
        for (size_t e = 0; e < Ne; ++e) {
            double X[8][3];
            for (int k = 0; k < 8; ++k)
                for (int d = 0; d < 3; ++d) X[k][d] = x[3 * (size_t)conn[8 * e + k] + d];
            double J[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double s = dN[i][0] * X[0][j];
                    for (int k = 1; k < 8; ++k) s += dN[i][k] * X[k][j];
                    J[i][j] = s;
                }
            double Ji[3][3];
            inverse3(J, Ji);
            const double det = det3(J);
            V0[e] = det * 8.0;
            for (int k = 0; k < 8; ++k)
                for (int i = 0; i < 3; ++i) {
                    double s = Ji[i][0] * dN[0][k];
                    s += Ji[i][1] * dN[1][k];
                    s += Ji[i][2] * dN[2][k];
                    BT[24 * e + 3 * k + i] = s;
                }
            const double nodeMass = rho * V0[e] / 8.0;
            for (int k = 0; k < 8; ++k) mass[conn[8 * e + k]] += nodeMass;
        }
        invM.resize(Nn);
        for (size_t i = 0; i < Nn; ++i) invM[i] = 1.0 / mass[i];
    }

    // Eigen determinant_impl<3> (Determinant.h:19-24, 56-65).
    static double det3(const double m[3][3]) {
        const double h0 = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]);
        const double h1 = m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]);
        const double h2 = m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
        return h0 - h1 + h2;
    }
    // Eigen compute_inverse<3> (InverseImpl.h:127-176).
    static inline double cof(const double m[3][3], int i, int j) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        return m[i1][j1] * m[i2][j2] - m[i1][j2] * m[i2][j1];
    }
    static void inverse3(const double m[3][3], double r[3][3]) {
        const double c00 = cof(m, 0, 0), c10 = cof(m, 1, 0), c20 = cof(m, 2, 0);
        const double det = c00 * m[0][0] + c10 * m[1][0] + c20 * m[2][0];
        const double invdet = 1.0 / det;
        r[1][0] = cof(m, 0, 1) * invdet;
        r[1][1] = cof(m, 1, 1) * invdet;
        r[2][0] = cof(m, 0, 2) * invdet;
        r[1][2] = cof(m, 2, 1) * invdet;
        r[2][1] = cof(m, 1, 2) * invdet;
        r[2][2] = cof(m, 2, 2) * invdet;
        r[0][0] = c00 * invdet;
        r[0][1] = c10 * invdet;
        r[0][2] = c20 * invdet;
    }

    // exDyna3D::min_ele_edge (exDyna3D.cpp:1173-1220).
    double min_edge(size_t e) const {
        static const int ed[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                      {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
        double mn = std::numeric_limits<double>::max();
        for (int q = 0; q < 12; ++q) {
            const double* p = &x[3 * (size_t)conn[8 * e + ed[q][0]]];
            const double* r = &x[3 * (size_t)conn[8 * e + ed[q][1]]];
            const double dx = r[0] - p[0], dy = r[1] - p[1], dz = r[2] - p[2];
            const double t = dx * dx + dy * dy + dz * dz;
            if (t < mn) mn = t;
        }
        return std::sqrt(mn);
    }

    // One element of add_inForce_to_rhs (exDyna3D.cpp:556-572): kinematics
    // (888-918, 838-885), Jaumann stress (920-998) [+ J2], hourglass
    // (1000-1071), bulk viscosity (1073-1109), stable dt (828-834).
    // Writes the element's nodal forces to fel; returns energy increments.
    void element(size_t e, double dt, double* en) {
        const uint32_t* cn = &conn[8 * e];
        double vel[3][8], xyz[8][3], ue[3][8];
        for (int k = 0; k < 8; ++k) {
            const size_t n = cn[k];
            for (int d = 0; d < 3; ++d) {
                vel[d][k] = v[3 * n + d];
                xyz[k][d] = x[3 * n + d];
                ue[d][k] = u[3 * n + d];
            }
        }
        const double* B = &BT[24 * e];  // B[3*k+j] = BT(k,j)
        // F = I + u_ele * BT (864)
        double F[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s = ue[i][0] * B[j];
                for (int k = 1; k < 8; ++k) s += ue[i][k] * B[3 * k + j];
                F[i][j] = (i == j ? 1.0 : 0.0) + s;
            }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Fel[9 * e + 3 * i + j] = F[i][j];
        const double J = det3(F);
        double Fi[3][3];
        inverse3(F, Fi);
        // bT = BT * F^-1 (874)
        double bT[8][3];
        for (int k = 0; k < 8; ++k)
            for (int j = 0; j < 3; ++j) {
                double s = B[3 * k + 0] * Fi[0][j];
                s += B[3 * k + 1] * Fi[1][j];
                s += B[3 * k + 2] * Fi[2][j];
                bT[k][j] = s;
            }
        // VolRate (884)
        double dots[3];
        for (int i = 0; i < 3; ++i) {
            double s = vel[i][0] * bT[0][i];
            for (int k = 1; k < 8; ++k) s += vel[i][k] * bT[k][i];
            dots[i] = s;
        }
        const double VolRate = dots[0] + dots[1] + dots[2];
        // L = vel * bT (928)
        double L[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s = vel[i][0] * bT[0][j];
                for (int k = 1; k < 8; ++k) s += vel[i][k] * bT[k][j];
                L[i][j] = s;
            }
        double eps[6];
        eps[0] = L[0][0];
        eps[1] = L[1][1];
        eps[2] = L[2][2];
        eps[3] = (L[0][1] + L[1][0]) / 2.0;
        eps[4] = (L[0][2] + L[2][0]) / 2.0;
        eps[5] = (L[1][2] + L[2][1]) / 2.0;
        for (int q = 0; q < 6; ++q) eps[q] *= dt;
        const double w0 = (L[0][1] - L[1][0]) / 2.0;
        const double w1 = (L[0][2] - L[2][0]) / 2.0;
        const double w2 = (L[1][2] - L[2][1]) / 2.0;
        const double tr = eps[0] + eps[1] + eps[2];
        const double twoG = 2 * G;
        double Jm[6];
        Jm[0] = lambda * tr + twoG * eps[0];
        Jm[1] = lambda * tr + twoG * eps[1];
        Jm[2] = lambda * tr + twoG * eps[2];
        Jm[3] = twoG * eps[3];
        Jm[4] = twoG * eps[4];
        Jm[5] = twoG * eps[5];
        double* s = &sigma[6 * e];
        double s_old[6];
        for (int q = 0; q < 6; ++q) s_old[q] = s[q];
        // Rotation terms as the reference writes them (964-970), including
        // its non-textbook shear terms.
        double inc[6];
        inc[0] = Jm[0] + 2 * s[3] * w0 * dt + 2 * s[4] * w1 * dt;
        inc[1] = Jm[1] - 2 * s[3] * w0 * dt + 2 * s[5] * w2 * dt;
        inc[2] = Jm[2] - 2 * s[4] * w1 * dt - 2 * s[5] * w2 * dt;
        inc[3] = Jm[3] + (s[1] - s[0]) * w0 * dt + s[5] * w1 * dt;
        inc[4] = Jm[4] + (s[2] - s[0]) * w1 * dt + s[5] * w0 * dt;
        inc[5] = Jm[5] + (s[2] - s[1]) * w2 * dt + s[3] * w1 * dt - s[4] * w0 * dt;
        for (int q = 0; q < 6; ++q) s[q] += inc[q];
        if (plastic) radial_return(s, epsp[e]);
        const double V0e = V0[e];
        const double cV = J * V0e;
        double f[8][3];
        for (int k = 0; k < 8; ++k) {
            const double fx = (bT[k][0] * s[0] + bT[k][1] * s[3] + bT[k][2] * s[4]) * cV;
            const double fy = (bT[k][0] * s[3] + bT[k][1] * s[1] + bT[k][2] * s[5]) * cV;
            const double fz = (bT[k][0] * s[4] + bT[k][1] * s[5] + bT[k][2] * s[2]) * cV;
            f[k][0] = 0.0 - fx;
            f[k][1] = 0.0 - fy;
            f[k][2] = 0.0 - fz;
        }
        // --- hourglass (1000-1071) ---
        static const double TAU[4][8] = {{1., -1., 1., -1., 1., -1., 1., -1.},
                                         {-1., -1., 1., 1., 1., 1., -1., -1.},
                                         {-1., 1., 1., -1., 1., -1., -1., 1.},
                                         {-1., 1., -1., 1., 1., -1., 1., -1.}};
        const double C2 = 0.1;
        const double R43 = 4.0 / 3.0;
        const double R1S8 = 1.0 / std::sqrt(8.0);
        double TX[4][3];
        for (int r = 0; r < 4; ++r)
            for (int j = 0; j < 3; ++j) {
                double t = TAU[r][0] * xyz[0][j];
                for (int k = 1; k < 8; ++k) t += TAU[r][k] * xyz[k][j];
                TX[r][j] = t;
            }
        double AT[4][8];
        for (int r = 0; r < 4; ++r)
            for (int k = 0; k < 8; ++k) {
                double t = TX[r][0] * B[3 * k + 0];
                t += TX[r][1] * B[3 * k + 1];
                t += TX[r][2] * B[3 * k + 2];
                AT[r][k] = TAU[r][k] - t;
            }
        double HIA[3][4];
        for (int i = 0; i < 3; ++i)
            for (int r = 0; r < 4; ++r) {
                double t = vel[i][0] * AT[r][0];
                for (int k = 1; k < 8; ++k) t += vel[i][k] * AT[r][k];
                HIA[i][r] = t;
            }
        double NX2 = 0.0;
        for (int q = 0; q < 24; ++q) NX2 += B[q] * B[q];
        const double V = J * V0e;
        const double DR = c_dr * (R43 * G + K) * C2 * NX2 * V * R1S8 * dt;
        const double CR = c_cr * rho * V0e / V * std::cbrt(V0e * V0e) * WOS / 4.0;
        double* Q = &qia[12 * e];
        for (int i = 0; i < 3; ++i)
            for (int r = 0; r < 4; ++r) Q[4 * i + r] += DR * HIA[i][r];
        double hgw = 0.0;
        for (int k = 0; k < 8; ++k)
            for (int i = 0; i < 3; ++i) {
                double hs = 0.0;
                for (int r = 0; r < 4; ++r) hs += Q[4 * i + r] * AT[r][k];
                double hd = AT[0][k] * HIA[i][0];
                for (int r = 1; r < 4; ++r) hd += AT[r][k] * HIA[i][r];
                hd *= CR;
                const double h = hs + hd;
                f[k][i] -= h;
                hgw += h * vel[i][k];
            }
        // --- bulk viscosity (1073-1109) ---
        double visw = 0.0;
        if (!(std::abs(VolRate) < volRateThreshold)) {
            const double Len = std::cbrt(V0e);
            const double rhoc = rho / J;
            const double p = rhoc * Len * (Cvisq * Cvisq * Len * VolRate * std::min(0.0, VolRate) - Cvisl * WOS * VolRate);
            for (int k = 0; k < 8; ++k)
                for (int i = 0; i < 3; ++i) f[k][i] += p * bT[k][i] * V;
            visw = -p * V * VolRate * dt;
        }
        // --- stable dt (828-834), L = min edge of the current x ---
        const double Le = Length[e];
        const double xi = Cvisl - Cvisq * Cvisq * Le / WOS * std::min(0.0, VolRate);
        dtStable[e] = (std::sqrt(1.0 + xi * xi) - xi) * Le / WOS;
        double* fo = &fel[24 * e];
        for (int k = 0; k < 8; ++k)
            for (int i = 0; i < 3; ++i) fo[3 * k + i] = f[k][i];
        // energies: stress work with the mid-step stress on the current volume
        double sw = 0.0;
        for (int q = 0; q < 3; ++q) sw += 0.5 * (s_old[q] + s[q]) * eps[q];
        for (int q = 3; q < 6; ++q) sw += (s_old[q] + s[q]) * eps[q];
        en[0] += sw * cV;
        en[1] += hgw * dt;
        en[2] += visw;
    }

    // J2 radial return, linear isotropic hardening (extension).
    void radial_return(double* s, double& ep) const {
        const double p = (s[0] + s[1] + s[2]) / 3.0;
        const double d0 = s[0] - p, d1 = s[1] - p, d2 = s[2] - p;
        const double j2 = 0.5 * (d0 * d0 + d1 * d1 + d2 * d2) + s[3] * s[3] + s[4] * s[4] + s[5] * s[5];
        const double q = std::sqrt(3.0 * j2);
        const double sy = yield0 + hard * ep;
        if (q > sy) {
            const double dg = (q - sy) / (3.0 * G + hard);
            const double ratio = 1.0 - 3.0 * G * dg / q;
            s[0] = d0 * ratio + p;
            s[1] = d1 * ratio + p;
            s[2] = d2 * ratio + p;
            s[3] *= ratio;
            s[4] *= ratio;
            s[5] *= ratio;
            ep += dg;
        }
    }

    // add_inForce_to_rhs (517-640) with the deterministic gather.
    void add_inForce_to_rhs(double dt) {
        double e0 = 0, e1 = 0, e2 = 0;
#pragma omp parallel num_threads(nthreads)
        {
            double en[3] = {0, 0, 0};
#pragma omp for schedule(static)
            for (long e = 0; e < (long)Ne; ++e) element((size_t)e, dt, en);
            const int t = omp_get_thread_num();
            thr_part[8 * t + 0] = en[0];
            thr_part[8 * t + 1] = en[1];
            thr_part[8 * t + 2] = en[2];
#pragma omp for schedule(static)
            for (long n = 0; n < (long)Nn; ++n) {
                double acc[3] = {0.0, 0.0, 0.0};
                double grp[3] = {0.0, 0.0, 0.0};
                int gk = -1;
                for (uint32_t q = csr_ptr[n]; q < csr_ptr[n + 1]; ++q) {
                    const uint32_t sl = csr_slot[q];
                    const int k = (int)(sl & 7u);
                    if (k != gk) {
                        if (gk >= 0)
                            for (int d = 0; d < 3; ++d) acc[d] += grp[d];
                        grp[0] = grp[1] = grp[2] = 0.0;
                        gk = k;
                    }
                    for (int d = 0; d < 3; ++d) grp[d] += fel[3 * (size_t)sl + d];
                }
                if (gk >= 0)
                    for (int d = 0; d < 3; ++d) acc[d] += grp[d];
                for (int d = 0; d < 3; ++d) a[3 * n + d] += acc[d];
            }
        }
        for (int t = 0; t < nthreads; ++t) {
            e0 += thr_part[8 * t + 0];
            e1 += thr_part[8 * t + 1];
            e2 += thr_part[8 * t + 2];
        }
        ie += e0;
        hg += e1;
        vis += e2;
    }

    void update_lengths() {
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (long e = 0; e < (long)Ne; ++e) Length[e] = min_edge((size_t)e);
    }

    double dts_min() const {
        double m = std::numeric_limits<double>::infinity();
        for (size_t e = 0; e < Ne; ++e) m = std::min(m, dtStable[e]);
        return m;
    }

    // apply_boundary_condition_vec (379-490): velocity lines, then SPC lines.
    // Energy account (extension): the kinetic energy the constraints add or
    // remove by setting v is external work, 1/2 m (v_new^2 - v_old^2) per
    // constrained dof, v_old being the value before this call (a dof named
    // by several lines counts once).
    std::vector<uint32_t> bc_stamp;
    uint32_t bc_tag = 0;
    std::vector<std::pair<size_t, double>> bc_old;
    void apply_bc_vec() {
        if (bc_stamp.size() != 3 * Nn) bc_stamp.assign(3 * Nn, 0);
        ++bc_tag;
        bc_old.clear();
        auto touch = [&](size_t idx) {
            if (bc_stamp[idx] != bc_tag) {
                bc_stamp[idx] = bc_tag;
                bc_old.push_back({idx, v[idx]});
            }
        };
        for (const ensc_bc_entry* b : cur_vel) {
            const double val = b->value * (b->amp >= 0 ? amp_interp(b->amp, time) : 1.0);
            for (uint32_t i = 0; i < b->n_nodes; ++i)
                for (int d = b->d0; d <= b->d1; ++d) {
                    touch(3 * (size_t)b->nodes[i] + d);
                    v[3 * (size_t)b->nodes[i] + d] = val;
                }
        }
        for (const ensc_bc_entry* b : cur_spc) {
            if (b->cond != 1) continue;
            for (uint32_t i = 0; i < b->n_nodes; ++i)
                for (int d = b->d0; d <= b->d1; ++d) {
                    touch(3 * (size_t)b->nodes[i] + d);
                    v[3 * (size_t)b->nodes[i] + d] = 0.0;
                }
        }
        double w = 0.0;
        for (const auto& p : bc_old) {
            const double vn = v[p.first];
            w += 0.5 * (vn * vn - p.second * p.second) / invM[p.first / 3];
        }
        wext += w;
    }

    // apply_boundary_condition_a (2029-2145) + the work of the reactions at
    // velocity-BC dofs, -(f_int + f_grav) v over the step.
    void apply_bc_a(double dt_avg) {
        std::fill(rf.begin(), rf.end(), 0.0);
        double w = 0.0;
        for (const ensc_bc_entry* b : cur_vel)
            for (uint32_t i = 0; i < b->n_nodes; ++i)
                for (int d = b->d0; d <= b->d1; ++d) {
                    const size_t idx = 3 * (size_t)b->nodes[i] + d;
                    w += -a[idx] * v[idx];
                    a[idx] = 0.0;
                }
        for (const ensc_bc_entry* b : cur_spc) {
            if (b->cond != 1) continue;
            for (uint32_t i = 0; i < b->n_nodes; ++i)
                for (int d = b->d0; d <= b->d1; ++d) {
                    const size_t idx = 3 * (size_t)b->nodes[i] + d;
                    const double sa = a[idx];
                    if (sa != 0.0) {
                        rf[idx] = -sa;
                        a[idx] = 0.0;
                    }
                }
        }
        wext += w * dt_avg;
    }

    // apply_gravity (800-826).  Remembers the step's amp * g and direction
    // for the work gravity does in nodal_update.
    bool g_now = false;
    double g_ag = 0.0, g_dir[3] = {0, 0, 0};
    void apply_gravity() {
        g_now = false;
        if (currentStepIndex >= grav.size()) return;
        const Gravity& g = grav[currentStepIndex];
        if (!g.on) return;
        const double amp = g.amp >= 0 ? deal_amp(g.amp) : 1.0;
        g_now = true;
        g_ag = amp * g.value;
        for (int d = 0; d < 3; ++d) g_dir[d] = g.dir[d];
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (long n = 0; n < (long)Nn; ++n)
            for (int d = 0; d < 3; ++d) a[3 * n + d] += amp * g.value * g.dir[d] / invM[n];
    }

    // Surface pressures of the current *Dsload lines (extension, see the
    // header): the faces of the step's lines (the step whose lines are
    // current: its own, else the previous step's unless op=NEW) in line
    // order; per node its faces in ascending face order.
    int32_t dsl_step = -1;
    std::vector<uint32_t> ds_nodes, ds_ptr, ds_list;
    std::vector<double> ds_val, ds_force, ds_fp;
    std::vector<int32_t> ds_amp;
    void set_dsload(size_t si) {
        if (!dsl[si].empty())
            dsl_step = (int32_t)si;
        else if (dsl_reset[si])
            dsl_step = -1;
        ds_nodes.clear();
        ds_val.clear();
        ds_amp.clear();
        if (dsl_step >= 0)
            for (const Dsload& L : dsl[(size_t)dsl_step])
                for (size_t q = 0; q < L.elems.size(); ++q) {
                    for (int c = 0; c < 4; ++c)
                        ds_nodes.push_back(conn[8 * (size_t)L.elems[q] + FACE_CORNERS[L.faces[q]][c]]);
                    ds_val.push_back(L.value);
                    ds_amp.push_back(L.amp);
                }
        ds_ptr.assign(Nn + 1, 0);
        for (uint32_t n : ds_nodes) ds_ptr[n + 1]++;
        for (size_t i = 0; i < Nn; ++i) ds_ptr[i + 1] += ds_ptr[i];
        ds_list.assign(ds_nodes.size(), 0);
        std::vector<uint32_t> fill(ds_ptr.begin(), ds_ptr.end() - 1);
        for (size_t i = 0; i < ds_nodes.size(); ++i) ds_list[fill[ds_nodes[i]]++] = (uint32_t)(i / 4);
        ds_force.assign(3 * ds_val.size(), 0.0);
        ds_fp.assign(ds_val.empty() ? 0 : 3 * Nn, 0.0);
    }

    // After apply_gravity: each face's force p amp A / 4, A = (x2 - x0) x
    // (x3 - x1) / 2 on the current positions, added to a per node in
    // ascending face order; ds_fp keeps the per-dof pressure force for the
    // work in nodal_update.
    void apply_dsload() {
        if (ds_val.empty()) return;
        for (size_t f = 0; f < ds_val.size(); ++f) {
            const double* p[4];
            for (int c = 0; c < 4; ++c) p[c] = &x[3 * (size_t)ds_nodes[4 * f + c]];
            const double e[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
            const double g[3] = {p[3][0] - p[1][0], p[3][1] - p[1][1], p[3][2] - p[1][2]};
            const double amp = ds_amp[f] >= 0 ? deal_amp(ds_amp[f]) : 1.0;
            const double s = ds_val[f] * amp * 0.125;
            ds_force[3 * f + 0] = s * (e[1] * g[2] - e[2] * g[1]);
            ds_force[3 * f + 1] = s * (e[2] * g[0] - e[0] * g[2]);
            ds_force[3 * f + 2] = s * (e[0] * g[1] - e[1] * g[0]);
        }
        for (size_t n = 0; n < Nn; ++n) {
            double fp[3] = {0.0, 0.0, 0.0};
            for (uint32_t q = ds_ptr[n]; q < ds_ptr[n + 1]; ++q)
                for (int d = 0; d < 3; ++d) {
                    a[3 * n + d] += ds_force[3 * (size_t)ds_list[q] + d];
                    fp[d] += ds_force[3 * (size_t)ds_list[q] + d];
                }
            for (int d = 0; d < 3; ++d) ds_fp[3 * n + d] = fp[d];
        }
    }

    // ---- fluid-structure coupling through virtual particles ----
    // get_fsi_sph_Ele_face_index (1282-1361): boundary nodes are held by
    // fewer than 8 elements; an element face is an FSI face when its 4
    // corners (Element_HexN8::Ele_face_indices, Element_HexN8.cpp:21-26) are
    // boundary nodes.  processInteractionElements (1584-1638) and
    // generateVirtualParticlesForFace (1428-1521) place nLayers particles per
    // FSI face; the map element -> particles (std::map, ascending element)
    // drives apply_fsiSph_nodeForce (1635-1689) and
    // update_virParticles_coor_vel (1691-1758).
    struct FsiElem {
        uint32_t e;
        std::vector<std::array<double, 3>> unit;  // unit coordinates of its particles
        std::vector<size_t> index;                // their global indices
    };
    std::vector<FsiElem> fsi;                     // ascending element
    std::vector<std::array<double, 3>> fsi_coord, fsi_vel, fsi_force;
    std::vector<double> fsi_fp;                   // [3Nn] the particles' nodal force of the step

    static double shape_value(int k, const double p[3]) {
        // Element_HexN8::get_shapeFunctionValue (Element_HexN8.cpp:46-93)
        static const double c[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                       {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
        return 1.0 / 8.0 * (1 + c[k][0] * p[0]) * (1 + c[k][1] * p[1]) * (1 + c[k][2] * p[2]);
    }
    static double shape_deriv(int k, int dir, const double p[3]) {
        // get_shapeFunctionDerivativeValue (Element_HexN8.cpp:103-161)
        static const double c[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                       {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
        if (dir == 0) return 1.0 / 8.0 * c[k][0] * (1 + c[k][1] * p[1]) * (1 + c[k][2] * p[2]);
        if (dir == 1) return 1.0 / 8.0 * (1 + c[k][0] * p[0]) * c[k][1] * (1 + c[k][2] * p[2]);
        return 1.0 / 8.0 * (1 + c[k][0] * p[0]) * (1 + c[k][1] * p[1]) * c[k][2];
    }
    // N(unit) . (nodal field of element e), sequential over the 8 corners
    void interp(uint32_t e, const double p[3], const std::vector<double>& f, double out[3]) const {
        for (int d = 0; d < 3; ++d) out[d] = 0.0;
        for (int k = 0; k < 8; ++k) {
            const double N = shape_value(k, p);
            for (int d = 0; d < 3; ++d) {
                const double t = N * f[3 * (size_t)conn[8 * (size_t)e + k] + d];
                out[d] = k ? out[d] + t : t;
            }
        }
    }

    size_t fsi_setup(int nLayers, double dist) {
        static const int faces[6][4] = {{3, 0, 4, 7}, {1, 2, 6, 5}, {0, 1, 5, 4},
                                        {2, 3, 7, 6}, {1, 0, 3, 2}, {4, 5, 6, 7}};
        static const double surf[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
        static const double nrm[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
        std::vector<int> held(Nn, 0);
        for (uint32_t n : conn) held[n]++;
        fsi.clear();
        fsi_coord.clear();
        fsi_vel.clear();
        for (uint32_t e = 0; e < Ne; ++e) {
            FsiElem E;
            E.e = e;
            for (int f = 0; f < 6; ++f) {
                int on = 0;
                for (int i = 0; i < 4; ++i) {
                    const int h = held[conn[8 * (size_t)e + faces[f][i]]];
                    if (h > 0 && h < 8) ++on;
                }
                if (on != 4) continue;
                double Jm[3][3];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) {
                        double s = 0.0;
                        for (int k = 0; k < 8; ++k) {
                            const double t = shape_deriv(k, i, surf[f]) * x[3 * (size_t)conn[8 * (size_t)e + k] + j];
                            s = k ? s + t : t;
                        }
                        Jm[i][j] = s;
                    }
                double nr[3];
                for (int i = 0; i < 3; ++i) nr[i] = Jm[i][0] * nrm[f][0] + Jm[i][1] * nrm[f][1] + Jm[i][2] * nrm[f][2];
                const double s2 = nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2];
                if (!(s2 > 1e-24)) continue;
                const double unit_step = dist / std::sqrt(s2);
                for (int k = 0; k < nLayers; ++k) {
                    std::array<double, 3> p;
                    for (int d = 0; d < 3; ++d) p[d] = surf[f][d] + k * unit_step * nrm[f][d];
                    double c[3], w[3];
                    interp(e, p.data(), x, c);
                    interp(e, p.data(), v, w);
                    E.unit.push_back(p);
                    E.index.push_back(fsi_coord.size());
                    fsi_coord.push_back({c[0], c[1], c[2]});
                    fsi_vel.push_back({w[0], w[1], w[2]});
                }
            }
            if (!E.unit.empty()) fsi.push_back(std::move(E));
        }
        fsi_force.assign(fsi_coord.size(), {0.0, 0.0, 0.0});
        fsi_fp.assign(fsi_coord.empty() ? 0 : 3 * Nn, 0.0);
        return fsi_coord.size();
    }

    // apply_fsiSph_nodeForce (1635-1689), on a = 0 before the internal forces
    void apply_fsi() {
        if (fsi_coord.empty()) return;
        for (const FsiElem& E : fsi) {
            double nf[8][3] = {};
            for (size_t q = 0; q < E.unit.size(); ++q) {
                const auto& F = fsi_force[E.index[q]];
                for (int k = 0; k < 8; ++k) {
                    const double N = shape_value(k, E.unit[q].data());
                    for (int d = 0; d < 3; ++d) nf[k][d] += F[d] * N;
                }
            }
            for (int k = 0; k < 8; ++k)
                for (int d = 0; d < 3; ++d) a[3 * (size_t)conn[8 * (size_t)E.e + k] + d] += nf[k][d];
        }
        fsi_fp = a;  // a held 0 before: the particles' nodal force
    }

    // update_virParticles_coor_vel (1691-1758), after move_mesh
    void update_particles() {
        for (const FsiElem& E : fsi)
            for (size_t q = 0; q < E.unit.size(); ++q) {
                double c[3], w[3];
                interp(E.e, E.unit[q].data(), x, c);
                interp(E.e, E.unit[q].data(), v, w);
                fsi_coord[E.index[q]] = {c[0], c[1], c[2]};
                fsi_vel[E.index[q]] = {w[0], w[1], w[2]};
            }
    }

    void nodal_update() {
        // compute_a, update_velocity, update_displacement, move_mesh (642-696)
        // Energy account (extension): gravity's and the pressures' work on
        // every dof over the step, f (v_old + v_new) / 2 dt_avg.
        double kep = 0.0, wg = 0.0, wp = 0.0, wf = 0.0;
        const double hs = dt_i + dt_i_1;
        const bool gon = g_now;
        const bool pon = !ds_val.empty();
        const bool fon = !fsi_coord.empty();
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : kep, wg, wp, wf)
        for (long n = 0; n < (long)Nn; ++n) {
            double k2 = 0.0;
            for (int d = 0; d < 3; ++d) {
                const size_t i = 3 * n + d;
                const double acc = a[i] * invM[n];
                const double vb = v[i];
                v[i] += hs * acc / 2.0;
                u[i] += dt_i * v[i];
                x[i] += dt_i * v[i];
                a[i] = 0.0;
                k2 += v[i] * v[i];
                if (gon) wg += g_ag * g_dir[d] / invM[n] * (vb + v[i]);
                if (pon) wp += ds_fp[i] * (vb + v[i]);
                if (fon) wf += fsi_fp[i] * (vb + v[i]);
            }
            kep += 0.5 * k2 / invM[n];
        }
        ke = kep;
        wext += 0.25 * hs * wg;
        wext += 0.25 * hs * wp;
        wext += 0.25 * hs * wf;
    }

    double kinetic() const {
        double k = 0.0;
        for (size_t n = 0; n < Nn; ++n) {
            double k2 = 0.0;
            for (int d = 0; d < 3; ++d) k2 += v[3 * n + d] * v[3 * n + d];
            k += 0.5 * k2 / invM[n];
        }
        return k;
    }

    // exDyna3D::init (105-142).
    void init() {
        setup();
        v = v_init;                 // apply_initial_condition (338-377)
        update_lengths();           // update_minVertex_perMaterial (656-661)
        // setCurrentStep(0) returns early: currentStepIndex is already 0.
        add_inForce_to_rhs(0.0);    // dt_i_1 = 0: only dtStable changes
        apply_gravity();
        apply_bc_a(0.0);            // no BCs are current yet
        std::fill(a.begin(), a.end(), 0.0);  // update_velocity with dt 0, then a = 0
        ie = hg = vis = wext = 0.0;
        ke = ke0 = kinetic();
    }

    // setCurrentStep (1760-1940).
    void setCurrentStep(size_t si) {
        if (si >= steps.size()) return;
        if (si == currentStepIndex) return;
        if (si > 0) {
            prev_spc = cur_spc;
            prev_vel = cur_vel;
        }
        currentStepIndex = si;
        const ensc_step_def& st = steps[si];
        if (si == 0) {
            cur_spc.clear();
            cur_vel.clear();
            for (uint32_t i = 0; i < st.n_spc; ++i) cur_spc.push_back(&st.spc[i]);
            for (uint32_t i = 0; i < st.n_vel; ++i) cur_vel.push_back(&st.vel[i]);
        } else {
            cur_spc.clear();
            for (uint32_t i = 0; i < st.n_spc; ++i) cur_spc.push_back(&st.spc[i]);
            if (!st.reset_spc) {
                if (!prev_spc.empty()) {
                    cur_spc.insert(cur_spc.end(), prev_spc.begin(), prev_spc.end());
                } else {
                    const ensc_step_def& ps = steps[si - 1];
                    for (uint32_t i = 0; i < ps.n_spc; ++i) cur_spc.push_back(&ps.spc[i]);
                }
            }
            cur_vel.clear();
            for (uint32_t i = 0; i < st.n_vel; ++i) cur_vel.push_back(&st.vel[i]);
            if (!st.reset_vel) cur_vel.insert(cur_vel.end(), prev_vel.begin(), prev_vel.end());
            if (!st.reset_dload) {
                if (!grav[si].on && grav[si - 1].on) grav[si] = grav[si - 1];
            }
            set_dsload(si);
        }
        apply_bc_vec();
    }

    // update_minVertex_perMaterial_and_dt (663-686).
    void update_dt() {
        update_lengths();
        double lmin = std::numeric_limits<double>::max();
        for (size_t e = 0; e < Ne; ++e)
            if (Length[e] < lmin) lmin = Length[e];
        last_lmin = lmin;
        dt_i_1 = dt_i;
        const double time_prev = time + dt_i_1;
        dt_i = lmin / WOS;
        const double tmp = dts_min();
        dt_i = dt_i < tmp ? dt_i : tmp;
        dt_i *= factor;
        if ((totalTime - time_prev) < dt_i && (totalTime - time_prev) > 0.0) {
            dt_i = totalTime - time_prev;
            dt_i_1 = dt_i;
        }
    }
    double last_lmin = 0.0;

    // output_results (207-336), called by computeSate after compute_a: at
    // an output time (or the total time) it records what the reference hands
    // to DataOut: node positions, a (= the summed force times invM,
    // compute_a 642-644), v and u, and per element the von Mises stress
    // (computeVonMises 1111-1133), the stress in VTK order and the
    // Green-Lagrange strain 0.5 (F^T F - I) of the pass's F (300-318).
    struct Snapshot {
        double time;
        std::vector<double> node;  // [Nn][12]: x, a, v, u
        std::vector<double> post;  // [Ne][13]: von Mises, stress (6), GL strain (6)
    };
    bool out_on = false;
    double time_output = 0.0, time_interval = 0.0;
    std::vector<Snapshot> snaps;
    void output_results() {
        if (!out_on) return;
        if (!(time >= time_output) && !(time >= totalTime)) return;
        time_output += time_interval;
        Snapshot S;
        S.time = time;
        S.node.resize(12 * Nn);
        for (size_t n = 0; n < Nn; ++n)
            for (int d = 0; d < 3; ++d) {
                S.node[12 * n + d] = x[3 * n + d];
                S.node[12 * n + 3 + d] = a[3 * n + d] * invM[n];
                S.node[12 * n + 6 + d] = v[3 * n + d];
                S.node[12 * n + 9 + d] = u[3 * n + d];
            }
        S.post.resize(13 * Ne);
        for (size_t e = 0; e < Ne; ++e) {
            const double* sg = &sigma[6 * e];
            const double t1 = (sg[0] - sg[1]) * (sg[0] - sg[1]);
            const double t2 = (sg[1] - sg[2]) * (sg[1] - sg[2]);
            const double t3 = (sg[2] - sg[0]) * (sg[2] - sg[0]);
            const double t4 = 6.0 * (sg[3] * sg[3] + sg[4] * sg[4] + sg[5] * sg[5]);
            double* o = &S.post[13 * e];
            o[0] = std::sqrt((t1 + t2 + t3 + t4) / 2.0);
            o[1] = sg[0];
            o[2] = sg[1];
            o[3] = sg[2];
            o[4] = sg[3];
            o[5] = sg[5];  // yz
            o[6] = sg[4];  // xz
            const double* F = &Fel[9 * e];
            double E[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    const double ftf = F[0 + i] * F[0 + j] + F[3 + i] * F[3 + j] + F[6 + i] * F[6 + j];
                    E[i][j] = 0.5 * (ftf - (i == j ? 1.0 : 0.0));
                }
            o[7] = E[0][0];
            o[8] = E[1][1];
            o[9] = E[2][2];
            o[10] = E[0][1];
            o[11] = E[1][2];
            o[12] = E[0][2];
        }
        snaps.push_back(std::move(S));
    }

    // computeSate (774-793).
    void computeSate() {
        time += dt_i_1;
        if (time > totalTime) time = totalTime;
        apply_bc_vec();
        apply_fsi();
        add_inForce_to_rhs(dt_i_1);
        apply_gravity();
        apply_dsload();
        apply_bc_a(0.5 * (dt_i + dt_i_1));
        output_results();
        nodal_update();
        update_particles();
        ++nsteps;
    }

    // run (54-102); the loop may be resumed after a max_steps stop.
    size_t run_step = 0;
    double step_end = 0.0;
    bool in_step = false;

    int64_t run(int64_t max_steps, const uint32_t* probe, uint32_t n_probe, double* hist, int64_t hist_rows) {
        int64_t taken = 0;
        if (run_step == 0) run_step = 1;
        while (run_step < steps.size()) {
            if (!in_step) {
                setCurrentStep(run_step);
                step_end = time + steps[run_step].period;
                in_step = true;
            }
            while (time < step_end) {
                if (max_steps > 0 && taken >= max_steps) return taken;
                update_dt();
                if (time + dt_i_1 > step_end) dt_i = dt_i_1 = step_end - time;
                computeSate();
                if (hist && taken < hist_rows) {
                    double* h = hist + taken * (ORACLE_HIST_BASE + 3 * (int64_t)n_probe);
                    h[OH_TIME] = time;
                    h[OH_DT_I] = dt_i;
                    h[OH_DT_I_1] = dt_i_1;
                    h[OH_DTS_MIN] = dts_min();
                    h[OH_KE] = ke;
                    h[OH_IE] = ie;
                    h[OH_HG] = hg;
                    h[OH_VIS] = vis;
                    h[OH_WEXT] = wext;
                    for (uint32_t p = 0; p < n_probe; ++p)
                        for (int d = 0; d < 3; ++d) h[ORACLE_HIST_BASE + 3 * p + d] = u[3 * (size_t)probe[p] + d];
                }
                ++taken;
            }
            in_step = false;
            ++run_step;
        }
        return taken;
    }
};

}  // namespace

extern "C" {

void* oracle_create(const ensc_model_desc* m, int n_threads) {
    Sim* s = new Sim();
    s->nthreads = n_threads > 0 ? n_threads : omp_get_max_threads();
    s->load(m);
    return s;
}

void oracle_destroy(void* h) { delete static_cast<Sim*>(h); }

int oracle_init(void* h) {
    static_cast<Sim*>(h)->init();
    return 0;
}

int64_t oracle_run(void* h, int64_t max_steps, const uint32_t* probe, uint32_t n_probe, double* hist, int64_t hist_rows) {
    return static_cast<Sim*>(h)->run(max_steps, probe, n_probe, hist, hist_rows);
}

void oracle_begin_step(void* h) {
    Sim* s = static_cast<Sim*>(h);
    if (s->run_step == 0) s->run_step = 1;
    if (!s->in_step && s->run_step < s->steps.size()) {
        s->setCurrentStep(s->run_step);
        s->step_end = s->time + s->steps[s->run_step].period;
        s->in_step = true;
    }
}

double oracle_time_steps(void* h, int64_t n_steps) {
    Sim* s = static_cast<Sim*>(h);
    oracle_begin_step(h);
    const auto t0 = std::chrono::steady_clock::now();
    for (int64_t i = 0; i < n_steps; ++i) {
        s->update_dt();
        if (s->time + s->dt_i_1 > s->step_end) s->dt_i = s->dt_i_1 = s->step_end - s->time;
        s->computeSate();
    }
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

int64_t oracle_get(void* h, const char* f, double* out) {
    Sim* s = static_cast<Sim*>(h);
    const std::vector<double>* src = nullptr;
    if (!std::strcmp(f, "u")) src = &s->u;
    else if (!std::strcmp(f, "v")) src = &s->v;
    else if (!std::strcmp(f, "x")) src = &s->x;
    else if (!std::strcmp(f, "a")) src = &s->a;
    else if (!std::strcmp(f, "rf")) src = &s->rf;
    else if (!std::strcmp(f, "sigma")) src = &s->sigma;
    else if (!std::strcmp(f, "qia")) src = &s->qia;
    else if (!std::strcmp(f, "epsp")) src = &s->epsp;
    else if (!std::strcmp(f, "dtStable")) src = &s->dtStable;
    else if (!std::strcmp(f, "Length")) src = &s->Length;
    else if (!std::strcmp(f, "V0")) src = &s->V0;
    else if (!std::strcmp(f, "B0")) src = &s->BT;
    else if (!std::strcmp(f, "invM")) src = &s->invM;
    if (!src) return -1;
    std::memcpy(out, src->data(), src->size() * sizeof(double));
    return (int64_t)src->size();
}

void oracle_scalars(void* h, double* o) {
    Sim* s = static_cast<Sim*>(h);
    o[OS_TIME] = s->time;
    o[OS_DT_I] = s->dt_i;
    o[OS_DT_I_1] = s->dt_i_1;
    o[OS_DTS_MIN] = s->dts_min();
    o[OS_KE] = s->ke;
    o[OS_IE] = s->ie;
    o[OS_HG] = s->hg;
    o[OS_VIS] = s->vis;
    o[OS_WEXT] = s->wext;
    o[OS_KE0] = s->ke0;
    o[OS_NSTEPS] = (double)s->nsteps;
    o[OS_STEP_INDEX] = (double)s->currentStepIndex;
    o[OS_LMIN] = s->last_lmin;
}

int oracle_threads(void* h) { return static_cast<Sim*>(h)->nthreads; }

void oracle_set_output(void* h, double interval) {
    Sim* s = static_cast<Sim*>(h);
    s->out_on = true;
    s->time_output = 0.0;  // init (exDyna3D.cpp:107)
    double ti = interval;
    if (!(ti > 0.0)) {
        // calculate_time_interval (1947-1977)
        const double T = s->totalTime;
        if (T <= 0.0) {
            ti = 1e-2;
        } else {
            ti = 1.0 / 50.0;
            const int frames = static_cast<int>(T / ti) + 1;
            if (frames < 10)
                ti = T / 10.0;
            else if (frames > 1000)
                ti = T / 1000.0;
        }
        if (ti > T && T > 0.0) ti = T;
    }
    s->time_interval = ti;
    s->snaps.clear();
}

int64_t oracle_fsi_setup(void* h, int n_layers, double dist) {
    return (int64_t)static_cast<Sim*>(h)->fsi_setup(n_layers, dist);
}

int oracle_fsi_set_forces(void* h, const double* f, int64_t n) {
    Sim* s = static_cast<Sim*>(h);
    if (n != (int64_t)s->fsi_force.size()) return -1;
    for (int64_t p = 0; p < n; ++p)
        for (int d = 0; d < 3; ++d) s->fsi_force[(size_t)p][d] = f[3 * p + d];
    return 0;
}

int64_t oracle_fsi_get(void* h, double* coords, double* vel) {
    Sim* s = static_cast<Sim*>(h);
    for (size_t p = 0; p < s->fsi_coord.size(); ++p)
        for (int d = 0; d < 3; ++d) {
            if (coords) coords[3 * p + d] = s->fsi_coord[p][d];
            if (vel) vel[3 * p + d] = s->fsi_vel[p][d];
        }
    return (int64_t)s->fsi_coord.size();
}

int64_t oracle_snapshots(void* h) { return (int64_t)static_cast<Sim*>(h)->snaps.size(); }

int64_t oracle_snapshot(void* h, int64_t k, const char* f, double* out) {
    Sim* s = static_cast<Sim*>(h);
    if (k < 0 || k >= (int64_t)s->snaps.size()) return -1;
    const Sim::Snapshot& S = s->snaps[(size_t)k];
    if (!std::strcmp(f, "time")) {
        out[0] = S.time;
        return 1;
    }
    static const char* node_f[4] = {"x", "a", "v", "u"};
    for (int q = 0; q < 4; ++q)
        if (!std::strcmp(f, node_f[q])) {
            for (size_t n = 0; n < s->Nn; ++n)
                for (int d = 0; d < 3; ++d) out[3 * n + d] = S.node[12 * n + 3 * q + d];
            return (int64_t)(3 * s->Nn);
        }
    int off = -1, nc = 0;
    if (!std::strcmp(f, "vonMises")) {
        off = 0;
        nc = 1;
    } else if (!std::strcmp(f, "Stress")) {
        off = 1;
        nc = 6;
    } else if (!std::strcmp(f, "Green-Lagrange Strain")) {
        off = 7;
        nc = 6;
    } else {
        return -1;
    }
    for (size_t e = 0; e < s->Ne; ++e)
        for (int q = 0; q < nc; ++q) out[(size_t)nc * e + q] = S.post[13 * e + off + q];
    return (int64_t)(nc * s->Ne);
}

}  // extern "C"
