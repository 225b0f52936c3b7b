// exDyna3D_mi355x.h - run exDyna3D's step loop on the MI355X engine.
//
// The binding a maintainer adds to the reference (ssadz/EnSCDynamics): a
// subclass of EnSC::exDyna3D (Explicit3D/include/exDyna3D.h:36-222) that
// hands the state DataIn filled to the engine's C ABI (include/ensc_engine.h,
// include/ensc_model.h), runs the steps on the GPU and copies the results
// back into the reference's members.  Two modes:
//   * whole run (default): run() (virtual, exDyna3D.h:46) executes every
//     *Step on the device in one call and then copies the state back;
//   * per step (stepwise = true): run() keeps the reference's loop shape
//     and calls computeSate() (virtual, exDyna3D.h:75) once per step
//     iteration; this class's computeSate() advances the engine by one step
//     (dt update, clamps and the whole of the reference's computeSate) and
//     updates time, dt_i and dt_i_1.  A further subclass that overrides
//     computeSate() around exDyna3D_MI355X::computeSate() - the reference's
//     own pattern for hooks - sees every step; pull_state() copies the
//     fields back into the reference's members when a hook needs them.
// FSI coupling: run() generates the virtual particles as the reference's
// run() does (get_fsiSph_virtualParticles_and_vel(0, 1/10),
// exDyna3D.cpp:67; fsi_layers_ and fsi_dist_ hold those arguments, so a
// coupler subclass may ask for particles), and setup_fsi() regenerates them
// on the current state.  The public fsi_share_data (exDyna3D.h:54) works as
// in the reference: the particle forces a coupler writes there are applied
// by the next step, and the particles' positions and velocities are copied
// back after every step (per-step mode) or after the run.
// The oracle's drop-in test (oracle/ref_harness/dropin.cpp) compiles this
// header against the reference's sources and runs both modes.
#pragma once

#include <sys/stat.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "ensc_engine.h"
#include "exDyna3D.h"

namespace EnSC {

class exDyna3D_MI355X : public exDyna3D {
public:
    // write_results: output_results' files (ResultData/solution_<n>.vtu,
    // exDyna3D.cpp:207-336) from device snapshots, as the reference writes
    // them.  stepwise: see above.
    explicit exDyna3D_MI355X(int precision = ENSC_FP64, bool write_results = true, bool stepwise = false)
        : precision_(precision), write_results_(write_results), stepwise_(stepwise) {}
    ~exDyna3D_MI355X() {
        if (eng_) ensc_engine_destroy(eng_);
    }

    // exDyna3D::run (exDyna3D.cpp:54-102) on the GPU.
    void run() override {
        // project.txt parameters and the solution vectors, as init() has them
        // (exDyna3D.cpp:105-118, 144-205)
        read_project_txt();
        init_data();
        mMatElastic.update();
        build_model();
        ensc_options o{};
        o.precision = precision_;
        o.reorder = 1;
        if (eng_) ensc_engine_destroy(eng_);
        eng_ = nullptr;
        check(ensc_engine_create(&m_, &o, &eng_));
        if (write_results_) {
            // the output interval as run() sets it (exDyna3D.cpp:65,
            // 1947-1977); the files go where DataOut::write_vtu puts them
            calculate_time_interval();
            ::mkdir("ResultData", 0755);
            check(ensc_engine_set_output(eng_, time_interval, 4, "ResultData"));
        }
        check(ensc_engine_init(eng_));
        setup_fsi(fsi_layers_, fsi_dist_);
        for (uint32_t s = 1; s < (uint32_t)steps.size(); ++s) {
            check(ensc_engine_set_step(eng_, s));
            step_done_ = 0;
            if (stepwise_) {
                while (!step_done_) computeSate();
            } else {
                push_fsi_forces();
                check(ensc_engine_advance(eng_, 0, &step_done_));
            }
        }
        pull_state();
    }

    long steps_run() const { return steps_run_; }

protected:
    // One step iteration on the GPU (the body of run()'s while loop,
    // exDyna3D.cpp:86-100, of which computeSate is the tail): only in the
    // stepwise mode, which run() drives.
    void computeSate() override {
        if (!eng_) throw std::runtime_error("EnSC-MI355X: computeSate outside run()");
        push_fsi_forces();
        int32_t done = 0;
        check(ensc_engine_advance(eng_, 1, &done));
        step_done_ = done;
        pull_scalars();
        pull_fsi_particles();
    }

    // get_fsiSph_virtualParticles_and_vel (exDyna3D.cpp:1363-1383) on the
    // engine: n_layers particles per FSI face, dist apart, on the current
    // state; fsi_share_data is sized to them, forces zero.
    void setup_fsi(int n_layers, double dist) {
        int64_t n = 0;
        check(ensc_engine_fsi_setup(eng_, n_layers, dist, &n));
        fsi_n_ = n;
        fsi_share_data.FSI_virtualParticles_coordinates.assign((size_t)n, {0.0, 0.0, 0.0});
        fsi_share_data.FSI_virtualParticles_velocity.assign((size_t)n, {0.0, 0.0, 0.0});
        fsi_share_data.FSI_virtualParticles_nodeForce.assign((size_t)n, {0.0, 0.0, 0.0});
        pull_fsi_particles();
    }

    void push_fsi_forces() {
        if (!fsi_n_) return;
        const auto& F = fsi_share_data.FSI_virtualParticles_nodeForce;
        if ((int64_t)F.size() != fsi_n_) throw std::runtime_error("EnSC-MI355X: FSI_virtualParticles_nodeForce resized");
        fsi_buf_.resize(3 * F.size());
        for (size_t p = 0; p < F.size(); ++p)
            for (int d = 0; d < 3; ++d) fsi_buf_[3 * p + d] = F[p][d];
        check(ensc_engine_fsi_set_forces(eng_, fsi_buf_.data(), fsi_n_));
    }

    void pull_fsi_particles() {
        if (!fsi_n_) return;
        std::vector<double> x(3 * (size_t)fsi_n_), v(3 * (size_t)fsi_n_);
        check(ensc_engine_fsi_get(eng_, x.data(), v.data(), fsi_n_));
        auto& X = fsi_share_data.FSI_virtualParticles_coordinates;
        auto& V = fsi_share_data.FSI_virtualParticles_velocity;
        for (size_t p = 0; p < (size_t)fsi_n_; ++p)
            for (int d = 0; d < 3; ++d) {
                X[p][d] = x[3 * p + d];
                V[p][d] = v[3 * p + d];
            }
    }

    // Copy the device state into the reference's members: solution_u, v, rf,
    // vertices, all_sigma, all_QIA, time, dt_i, dt_i_1.
    void pull_state() {
        const size_t Nn = vertices.size(), Ne = hexahedron_elements.size();
        std::vector<double> buf(3 * Nn);
        check(ensc_engine_download(eng_, "u", solution_u.data(), (int64_t)(3 * Nn)));
        check(ensc_engine_download(eng_, "v", solution_v.data(), (int64_t)(3 * Nn)));
        check(ensc_engine_download(eng_, "rf", solution_rf.data(), (int64_t)(3 * Nn)));
        check(ensc_engine_download(eng_, "x", buf.data(), (int64_t)buf.size()));
        for (size_t i = 0; i < Nn; ++i)
            for (int d = 0; d < 3; ++d) vertices[i][d] = buf[3 * i + d];
        pull_fsi_particles();
        std::vector<double> s(6 * Ne), qv(12 * Ne);
        check(ensc_engine_download(eng_, "sigma", s.data(), (int64_t)s.size()));
        check(ensc_engine_download(eng_, "qia", qv.data(), (int64_t)qv.size()));
        for (size_t e = 0; e < Ne; ++e) {
            for (int k = 0; k < 6; ++k) all_sigma[e][k] = s[6 * e + k];
            for (int k = 0; k < 12; ++k) all_QIA[e][k] = qv[12 * e + k];
        }
        pull_scalars();
    }

    void pull_scalars() {
        double sc[ENSC_S_COUNT];
        check(ensc_engine_scalars(eng_, sc));
        time = sc[ENSC_S_TIME];
        dt_i = sc[ENSC_S_DT_I];
        dt_i_1 = sc[ENSC_S_DT_I_1];
        steps_run_ = (long)sc[ENSC_S_NSTEPS];
    }

    ensc_engine* engine() const { return eng_; }

    int precision_;
    bool write_results_, stepwise_;
    long steps_run_ = 0;
    int32_t step_done_ = 0;
    ensc_engine* eng_ = nullptr;
    // the arguments of run()'s particle generation (exDyna3D.cpp:67)
    int fsi_layers_ = 0;
    double fsi_dist_ = 1.0 / 10.0;
    int64_t fsi_n_ = 0;
    std::vector<double> fsi_buf_;

private:
    // The state DataIn filled (exDyna3D.h:141-220), flattened into the C
    // model description; the arrays it points at are kept in members.
    void build_model() {
        const size_t Nn = vertices.size(), Ne = hexahedron_elements.size();
        x0.assign(3 * Nn, 0.0);
        v0.assign(3 * Nn, 0.0);
        conn.assign(8 * Ne, 0);
        for (size_t i = 0; i < Nn; ++i)
            for (int d = 0; d < 3; ++d) x0[3 * i + d] = vertices[i][d];
        for (size_t e = 0; e < Ne; ++e)
            for (int k = 0; k < 8; ++k) conn[8 * e + k] = hexahedron_elements[e].get_verticesIndex()[k];
        auto nodes_of = [&](const std::string& set) {
            std::vector<uint32_t> ids;
            if (!set.empty() && set[0] == '#')
                ids.push_back((uint32_t)(std::stoi(set.substr(1)) - 1));
            else if (map_set_node_list.count(set))
                for (size_t n : map_set_node_list.at(set)) ids.push_back((uint32_t)n);
            return ids;
        };
        for (const auto& iv : ini_vel_generation)  // apply_initial_condition (338-377)
            for (uint32_t n : nodes_of(iv.first)) v0[3 * n + iv.second.first] = iv.second.second;
        // amplitudes by name (map_amp_list, exDyna3D.h:194)
        amp_names.clear();
        at.clear();
        aa.clear();
        for (const auto& kv : map_amp_list) {
            amp_names.push_back(kv.first);
            at.emplace_back();
            aa.emplace_back();
            for (const auto& p : kv.second) {
                at.back().push_back(p[0]);
                aa.back().push_back(p[1]);
            }
        }
        amps.clear();
        for (size_t i = 0; i < at.size(); ++i) amps.push_back({at[i].data(), aa[i].data(), (uint32_t)at[i].size(), 0});
        auto amp_id = [&](const std::string& n) {
            for (size_t i = 0; i < amp_names.size(); ++i)
                if (amp_names[i] == n) return (int32_t)i;
            return (int32_t)-1;
        };
        // steps (StepData.h:14-34) with their BC lines (BoundaryCondition.h:13-19)
        keep.assign(steps.size(), {});
        spc.assign(steps.size(), {});
        vel.assign(steps.size(), {});
        sd.assign(steps.size(), ensc_step_def{});
        for (size_t s = 0; s < steps.size(); ++s) {
            const StepData& st = steps[s];
            for (const auto& b : st.boundary.spc_nodes) {
                keep[s].push_back(nodes_of(b.first));
                spc[s].push_back({nullptr, 0, (int32_t)b.second[0], (int32_t)b.second[1], (int32_t)b.second[2], 0.0, -1, 0});
            }
            for (const auto& b : st.boundary.vel_nodes) {
                keep[s].push_back(nodes_of(b.first));
                vel[s].push_back({nullptr, 0, (int32_t)b.second.first[0], (int32_t)b.second.first[1], 1, b.second.second, -1, 0});
            }
            size_t q = 0;
            for (auto& e : spc[s]) {
                e.nodes = keep[s][q].data();
                e.n_nodes = (uint32_t)keep[s][q].size();
                ++q;
            }
            for (auto& e : vel[s]) {
                e.nodes = keep[s][q].data();
                e.n_nodes = (uint32_t)keep[s][q].size();
                ++q;
            }
            const auto& g = st.gravity;
            sd[s] = ensc_step_def{};
            sd[s].period = st.timePeriod;
            sd[s].reset_spc = st.resetSpcBoundary;
            sd[s].reset_vel = st.resetVelBoundary;
            sd[s].reset_dload = st.resetDload;
            sd[s].grav_on = std::get<0>(g);
            sd[s].grav_amp = amp_id(std::get<1>(g));
            sd[s].grav_value = std::get<2>(g);
            sd[s].grav_dir[0] = std::get<3>(g);
            sd[s].grav_dir[1] = std::get<4>(g);
            sd[s].grav_dir[2] = std::get<5>(g);
            sd[s].n_spc = (uint32_t)spc[s].size();
            sd[s].spc = spc[s].data();
            sd[s].n_vel = (uint32_t)vel[s].size();
            sd[s].vel = vel[s].data();
        }
        m_ = ensc_model_desc{};
        ensc_model_desc& m = m_;
        m.n_nodes = (uint32_t)Nn;
        m.n_elems = (uint32_t)Ne;
        m.x0 = x0.data();
        m.conn = conn.data();
        m.E = mMatElastic.E;
        m.nu = mMatElastic.v;
        m.rho = mMatElastic.rho;
        m.c_dr = c_dr;
        m.c_cr = c_cr;
        m.factor = factor_timeStep;
        m.cvisl = Cvisl;
        m.cvisq = Cvisq;
        m.total_time = totalTime;
        m.v_init = v0.data();
        m.n_steps = (uint32_t)sd.size();
        m.steps = sd.data();
        m.n_amps = (uint32_t)amps.size();
        m.amps = amps.empty() ? nullptr : amps.data();
    }

    static void check(ensc_status s) {
        if (s != ENSC_OK) throw std::runtime_error(std::string("EnSC-MI355X: ") + ensc_last_error());
    }

    // storage behind m_
    ensc_model_desc m_{};
    std::vector<double> x0, v0;
    std::vector<uint32_t> conn;
    std::vector<std::string> amp_names;
    std::vector<std::vector<double>> at, aa;
    std::vector<ensc_amplitude> amps;
    std::vector<std::vector<std::vector<uint32_t>>> keep;
    std::vector<std::vector<ensc_bc_entry>> spc, vel;
    std::vector<ensc_step_def> sd;
};

}  // namespace EnSC
