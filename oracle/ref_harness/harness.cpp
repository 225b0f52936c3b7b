// Test infrastructure only: a driver that runs the reference solver
// (ssadz/EnSCDynamics, compiled from its own sources by oracle/Makefile)
// on one .inp deck and writes what it computes as golden vectors.
//
// Nothing in the engine links or loads this.  It exists so the oracle
// restatement and the reader can be checked against the real reference.
//
// The reference writes VTU files through VTK, which is not available here,
// so DataOut (Explicit3D/include/dataOut.h:44-97) is given empty bodies
// (dataout_stub.cpp).
// Stepping is observed by subclassing exDyna3D and overriding the virtual
// computeSate() (Explicit3D/include/exDyna3D.h:75): the base step runs
// unchanged, then the state is recorded.
//
// With ENSC_FSI=layers,dist,scale in the environment the subclass also acts
// as an FSI coupler: before the first step it generates the virtual
// particles with the reference's own get_fsiSph_virtualParticles_and_vel
// (exDyna3D.cpp:1363-1383) and, before every step n, writes the force
// scale * sin(0.37 (p + 1) + 1.1 d + 0.05 n) on particle p, component d,
// into fsi_share_data (exDyna3D.h:54), which the base step applies.
#include "exDyna3D.h"
#include "dataIn.h"
#include "dataOut.h"
#include "spdlog/spdlog.h"

#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <unistd.h>

#include "npy.h"

using npy::npy_f64;
using npy::npy_u32;

namespace {

std::string jstr(const std::string& s) {
    std::string o = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o + "\"";
}

std::string jnum(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%.17g", v);
    return b;
}

struct StopRun {};

}  // namespace

class Probe : public EnSC::exDyna3D {
public:
    std::string out;
    long max_steps = -1;
    std::vector<long> snap_steps;
    std::vector<uint32_t> probe_nodes;
    long nstep = 0;
    std::vector<double> hist;  // per step: time, dt_i, dt_i_1, min dtStable, then 3 u per probe node

    size_t nn() const { return vertices.size(); }

    void dump_model() {
        const size_t Nn = vertices.size(), Ne = hexahedron_elements.size();
        std::vector<double> x0(3 * Nn);
        for (size_t i = 0; i < Nn; ++i)
            for (int d = 0; d < 3; ++d) x0[3 * i + d] = vertices[i][d];
        npy_f64(out + "/x0.npy", {Nn, 3}, x0.data());
        std::vector<uint32_t> conn(8 * Ne);
        for (size_t e = 0; e < Ne; ++e) {
            const auto& vi = hexahedron_elements[e].get_verticesIndex();
            for (int k = 0; k < 8; ++k) conn[8 * e + k] = vi[k];
        }
        npy_u32(out + "/conn.npy", {Ne, 8}, conn.data());

        std::ofstream j(out + "/model.json");
        j << "{\n";
        j << "\"E\": " << jnum(mMatElastic.E) << ", \"nu\": " << jnum(mMatElastic.v) << ", \"rho\": " << jnum(mMatElastic.rho) << ",\n";
        j << "\"G\": " << jnum(mMatElastic.G) << ", \"K\": " << jnum(mMatElastic.K) << ", \"lambda\": " << jnum(mMatElastic.lambda)
          << ", \"WOS\": " << jnum(mMatElastic.WOS) << ",\n";
        j << "\"totalTime\": " << jnum(totalTime) << ", \"Cvisl\": " << jnum(Cvisl) << ", \"Cvisq\": " << jnum(Cvisq) << ",\n";
        j << "\"c_dr\": " << jnum(c_dr) << ", \"c_cr\": " << jnum(c_cr) << ", \"factor\": " << jnum(factor_timeStep) << ",\n";
        j << "\"n_nodes\": " << Nn << ", \"n_elems\": " << Ne << ",\n";
        j << "\"ini_vel\": [";
        for (size_t i = 0; i < ini_vel_generation.size(); ++i) {
            const auto& iv = ini_vel_generation[i];
            j << (i ? ", " : "") << "[" << jstr(iv.first) << ", " << iv.second.first << ", " << jnum(iv.second.second) << "]";
        }
        j << "],\n\"node_sets\": {";
        bool first = true;
        for (const auto& kv : map_set_node_list) {
            j << (first ? "" : ", ") << jstr(kv.first) << ": [";
            for (size_t i = 0; i < kv.second.size(); ++i) j << (i ? ", " : "") << kv.second[i];
            j << "]";
            first = false;
        }
        j << "},\n\"amplitudes\": {";
        first = true;
        for (const auto& kv : map_amp_list) {
            j << (first ? "" : ", ") << jstr(kv.first) << ": [";
            for (size_t i = 0; i < kv.second.size(); ++i)
                j << (i ? ", " : "") << "[" << jnum(kv.second[i][0]) << ", " << jnum(kv.second[i][1]) << "]";
            j << "]";
            first = false;
        }
        j << "},\n\"steps\": [\n";
        for (size_t s = 0; s < steps.size(); ++s) {
            const auto& st = steps[s];
            j << (s ? ",\n" : "") << "{\"name\": " << jstr(st.name) << ", \"period\": " << jnum(st.timePeriod)
              << ", \"reset_spc\": " << (st.resetSpcBoundary ? 1 : 0) << ", \"reset_vel\": " << (st.resetVelBoundary ? 1 : 0)
              << ", \"reset_dload\": " << (st.resetDload ? 1 : 0) << ", \"reset_dsload\": " << (st.resetDsload ? 1 : 0);
            j << ", \"spc\": [";
            for (size_t i = 0; i < st.boundary.spc_nodes.size(); ++i) {
                const auto& b = st.boundary.spc_nodes[i];
                j << (i ? ", " : "") << "[" << jstr(b.first) << ", " << b.second[0] << ", " << b.second[1] << ", " << b.second[2] << "]";
            }
            j << "], \"vel\": [";
            for (size_t i = 0; i < st.boundary.vel_nodes.size(); ++i) {
                const auto& b = st.boundary.vel_nodes[i];
                j << (i ? ", " : "") << "[" << jstr(b.first) << ", " << b.second.first[0] << ", " << b.second.first[1] << ", "
                  << jnum(b.second.second) << "]";
            }
            const auto& g = st.gravity;
            j << "], \"gravity\": [" << (std::get<0>(g) ? 1 : 0) << ", " << jstr(std::get<1>(g)) << ", " << jnum(std::get<2>(g)) << ", "
              << jnum(std::get<3>(g)) << ", " << jnum(std::get<4>(g)) << ", " << jnum(std::get<5>(g)) << "]";
            j << ", \"n_dsload\": " << st.dsload.size() << "}";
        }
        j << "\n]\n}\n";
    }

    void dump_setup() {
        // State right after init(): B0, V0, inverse lumped mass, the init
        // pass's dtStable and min edge lengths.
        const size_t Nn = vertices.size(), Ne = hexahedron_elements.size();
        std::vector<double> b0(24 * Ne);
        for (size_t e = 0; e < Ne; ++e)
            for (int a = 0; a < 8; ++a)
                for (int d = 0; d < 3; ++d) b0[24 * e + 3 * a + d] = all_BT[e](a, d);
        npy_f64(out + "/B0.npy", {Ne, 8, 3}, b0.data());
        npy_f64(out + "/V0.npy", {Ne}, all_Volume.data());
        std::vector<double> im(Nn);
        for (size_t i = 0; i < Nn; ++i) im[i] = inv_mass_matrix[3 * i];
        npy_f64(out + "/invM.npy", {Nn}, im.data());
        npy_f64(out + "/dtStable0.npy", {Ne}, dtStable.data());
        npy_f64(out + "/Length0.npy", {Ne}, all_Length.data());
        npy_f64(out + "/v_init.npy", {Nn, 3}, solution_v.data());
    }

    void dump_state(const std::string& tag) {
        const size_t Nn = vertices.size(), Ne = hexahedron_elements.size();
        npy_f64(out + "/" + tag + "_u.npy", {Nn, 3}, solution_u.data());
        npy_f64(out + "/" + tag + "_v.npy", {Nn, 3}, solution_v.data());
        npy_f64(out + "/" + tag + "_rf.npy", {Nn, 3}, solution_rf.data());
        std::vector<double> x(3 * Nn);
        for (size_t i = 0; i < Nn; ++i)
            for (int d = 0; d < 3; ++d) x[3 * i + d] = vertices[i][d];
        npy_f64(out + "/" + tag + "_x.npy", {Nn, 3}, x.data());
        std::vector<double> s(6 * Ne), q(12 * Ne);
        for (size_t e = 0; e < Ne; ++e) {
            for (int k = 0; k < 6; ++k) s[6 * e + k] = all_sigma[e][k];
            for (int k = 0; k < 12; ++k) q[12 * e + k] = all_QIA[e][k];
        }
        npy_f64(out + "/" + tag + "_sigma.npy", {Ne, 6}, s.data());
        npy_f64(out + "/" + tag + "_qia.npy", {Ne, 12}, q.data());
        npy_f64(out + "/" + tag + "_dtStable.npy", {Ne}, dtStable.data());
        npy_f64(out + "/" + tag + "_Length.npy", {Ne}, all_Length.data());
        double sc[5] = {time, dt_i, dt_i_1, (double)nstep, (double)currentStepIndex};
        npy_f64(out + "/" + tag + "_scalars.npy", {5}, sc);
    }

    void record() {
        hist.push_back(time);
        hist.push_back(dt_i);
        hist.push_back(dt_i_1);
        hist.push_back(dtStable.minCoeff());
        for (uint32_t n : probe_nodes)
            for (int d = 0; d < 3; ++d) hist.push_back(solution_u[3 * (size_t)n + d]);
    }

    int fsi_layers = 0;
    double fsi_dist = 0.0, fsi_scale = 0.0;

    void dump_particles(const std::string& tag) {
        const auto& X = fsi_share_data.FSI_virtualParticles_coordinates;
        const auto& V = fsi_share_data.FSI_virtualParticles_velocity;
        std::vector<double> x(3 * X.size()), v(3 * V.size());
        for (size_t p = 0; p < X.size(); ++p)
            for (int d = 0; d < 3; ++d) {
                x[3 * p + d] = X[p][d];
                v[3 * p + d] = V[p][d];
            }
        npy_f64(out + "/" + tag + "_fsi_x.npy", {X.size(), 3}, x.data());
        npy_f64(out + "/" + tag + "_fsi_v.npy", {V.size(), 3}, v.data());
    }

    void computeSate() override {
        if (fsi_layers > 0) {
            if (nstep == 0) {
                get_fsiSph_virtualParticles_and_vel(fsi_layers, fsi_dist);
                dump_particles("setup");
            }
            auto& F = fsi_share_data.FSI_virtualParticles_nodeForce;
            for (size_t p = 0; p < F.size(); ++p)
                for (int d = 0; d < 3; ++d) F[p][d] = fsi_scale * std::sin(0.37 * (double)(p + 1) + 1.1 * d + 0.05 * (double)nstep);
        }
        EnSC::exDyna3D::computeSate();
        ++nstep;
        record();
        for (long s : snap_steps)
            if (s == nstep) dump_state("step" + std::to_string(nstep));
        if (max_steps > 0 && nstep >= max_steps) throw StopRun();
    }

    void finish() {
        size_t w = 4 + 3 * probe_nodes.size();
        npy_f64(out + "/history.npy", {hist.size() / w, w}, hist.data());
        dump_state("final");
        if (fsi_layers > 0) dump_particles("final");
    }

    void do_init() { init(); }
};

static std::vector<long> parse_list(const char* s) {
    std::vector<long> v;
    std::stringstream ss(s);
    std::string t;
    while (std::getline(ss, t, ','))
        if (!t.empty()) v.push_back(std::stol(t));
    return v;
}

int main(int argc, char** argv) {
    // usage: ref_harness WORKDIR DECK OUTDIR [max_steps] [snap,steps] [probe,nodes] [c_dr c_cr factor]
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s WORKDIR DECK OUTDIR [max_steps] [snaps] [probes] [c_dr c_cr factor]\n", argv[0]);
        return 2;
    }
    std::string work = argv[1], deck = argv[2], out = argv[3];
    spdlog::set_level(spdlog::level::warn);
    if (chdir(work.c_str()) != 0) {
        std::perror("chdir");
        return 2;
    }
    {
        // project.txt is read from the working directory (exDyna3D.cpp:705-712).
        std::ofstream p("project.txt");
        p << "#deck\n" << deck << "\n";
        if (argc >= 10) {
            p << "#\n#c_dr\n" << argv[7] << "\n#\n#c_cr\n" << argv[8] << "\n#\n#factor\n" << argv[9] << "\n#\n#time_interval\n1.0\n";
        }
    }
    Probe sim;
    sim.out = out;
    if (argc >= 5) sim.max_steps = std::atol(argv[4]);
    if (argc >= 6) sim.snap_steps = parse_list(argv[5]);
    if (argc >= 7)
        for (long n : parse_list(argv[6])) sim.probe_nodes.push_back((uint32_t)n);
    if (const char* f = std::getenv("ENSC_FSI")) {
        if (std::sscanf(f, "%d,%lf,%lf", &sim.fsi_layers, &sim.fsi_dist, &sim.fsi_scale) != 3) {
            std::fprintf(stderr, "ENSC_FSI must be layers,dist,scale\n");
            return 2;
        }
    }
    EnSC::DataIn reader(sim);
    reader.read_inp(deck);
    sim.do_init();
    // After init(): read_project_txt has set c_dr, c_cr and factor_timeStep
    // (exDyna3D.cpp:698-772); node coordinates have not moved yet.
    sim.dump_model();
    sim.dump_setup();
    try {
        sim.run();
    } catch (const StopRun&) {
    }
    sim.finish();
    std::printf("steps %ld\n", sim.nstep);
    return 0;
}
