// Test infrastructure only: the reference solver with its step loop on the
// MI355X engine.  The reference's own reader (DataIn, dataIn.cpp:300-558)
// fills an exDyna3D_MI355X (enscdynamics_amd/binding/exDyna3D_mi355x.h, the
// subclass a maintainer adds to the reference), whose run() executes on the
// GPU through the C ABI; the final state is written for comparison with the
// CPU reference's goldens.  With "stepwise" the binding's per-step mode runs,
// and this subclass hooks computeSate() the way the reference's own
// subclasses do, counting the step iterations it sees.  With
// ENSC_FSI=layers,dist,scale it is also an FSI coupler, as the reference
// harness's subclass (ref_harness/harness.cpp): particles generated at the
// first step, force scale * sin(0.37 (p + 1) + 1.1 d + 0.05 n) written into
// fsi_share_data before step n; the particles are written at the end.
//
// usage: ref_dropin WORKDIR DECK OUTDIR fp64|mixed [c_dr c_cr factor [stepwise]]
#include "dataIn.h"
#include "exDyna3D_mi355x.h"
#include "npy.h"
#include "spdlog/spdlog.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <unistd.h>

namespace {
class DropIn : public EnSC::exDyna3D_MI355X {
public:
    using exDyna3D_MI355X::exDyna3D_MI355X;
    long hooked = 0;  // computeSate() calls seen by this subclass
    int fsi_layers = 0;
    double fsi_dist = 0.0, fsi_scale = 0.0;
    void computeSate() override {
        if (fsi_layers > 0) {
            if (hooked == 0) setup_fsi(fsi_layers, fsi_dist);
            auto& F = fsi_share_data.FSI_virtualParticles_nodeForce;
            for (size_t p = 0; p < F.size(); ++p)
                for (int d = 0; d < 3; ++d) F[p][d] = fsi_scale * std::sin(0.37 * (double)(p + 1) + 1.1 * d + 0.05 * (double)hooked);
        }
        exDyna3D_MI355X::computeSate();
        ++hooked;
    }
    void write(const std::string& out) const {
        const size_t Nn = vertices.size(), Ne = hexahedron_elements.size();
        std::vector<double> u(3 * Nn), x(3 * Nn), s(6 * Ne);
        for (size_t i = 0; i < 3 * Nn; ++i) u[i] = solution_u[i];
        for (size_t i = 0; i < Nn; ++i)
            for (int d = 0; d < 3; ++d) x[3 * i + d] = vertices[i][d];
        for (size_t e = 0; e < Ne; ++e)
            for (int k = 0; k < 6; ++k) s[6 * e + k] = all_sigma[e][k];
        npy::npy_f64(out + "/u.npy", {Nn, 3}, u.data());
        npy::npy_f64(out + "/x.npy", {Nn, 3}, x.data());
        npy::npy_f64(out + "/sigma.npy", {Ne, 6}, s.data());
        if (fsi_layers > 0) {
            const auto& X = fsi_share_data.FSI_virtualParticles_coordinates;
            const auto& V = fsi_share_data.FSI_virtualParticles_velocity;
            std::vector<double> px(3 * X.size()), pv(3 * V.size());
            for (size_t p = 0; p < X.size(); ++p)
                for (int d = 0; d < 3; ++d) {
                    px[3 * p + d] = X[p][d];
                    pv[3 * p + d] = V[p][d];
                }
            npy::npy_f64(out + "/fsi_x.npy", {X.size(), 3}, px.data());
            npy::npy_f64(out + "/fsi_v.npy", {V.size(), 3}, pv.data());
        }
        std::ofstream f(out + "/scalars.txt");
        f.precision(17);
        f << "time " << time << "\nsteps " << steps_run() << "\nhooked " << hooked << "\n";
    }
};
}  // namespace

int main(int argc, char** argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s WORKDIR DECK OUTDIR fp64|mixed [c_dr c_cr factor [stepwise]]\n", argv[0]);
        return 2;
    }
    const std::string work = argv[1], deck = argv[2], out = argv[3], prec = argv[4];
    spdlog::set_level(spdlog::level::warn);
    if (chdir(work.c_str()) != 0) {
        std::perror("chdir");
        return 2;
    }
    {
        // project.txt is read from the working directory (exDyna3D.cpp:705-712)
        std::ofstream p("project.txt");
        p << "#deck\n" << deck << "\n";
        if (argc >= 8)
            p << "#\n#c_dr\n" << argv[5] << "\n#\n#c_cr\n" << argv[6] << "\n#\n#factor\n" << argv[7]
              << "\n#\n#time_interval\n1.0\n";
    }
    const bool stepwise = argc >= 9 && std::string(argv[8]) == "stepwise";
    DropIn sim(prec == "mixed" ? ENSC_MIXED : ENSC_FP64, true, stepwise);
    if (const char* f = std::getenv("ENSC_FSI")) {
        if (std::sscanf(f, "%d,%lf,%lf", &sim.fsi_layers, &sim.fsi_dist, &sim.fsi_scale) != 3 || !stepwise) {
            std::fprintf(stderr, "ENSC_FSI must be layers,dist,scale, with stepwise\n");
            return 2;
        }
    }
    EnSC::DataIn reader(sim);
    reader.read_inp(deck);
    try {
        sim.run();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    sim.write(out);
    std::printf("steps %ld\n", sim.steps_run());
    return 0;
}
