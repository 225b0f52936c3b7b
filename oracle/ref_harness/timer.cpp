// Test infrastructure only: times the reference solver's own step loop
// (ssadz/EnSCDynamics, compiled from its sources by oracle/Makefile) on one
// deck, as the CPU baseline beside the GPU engine.
//
// usage: ref_timer WORKDIR DECK WARM STEPS [c_dr c_cr factor]
//
// The deck is read by the reference's DataIn and initialised by its init();
// then exDyna3D::run executes unchanged.  The timer starts when computeSate
// (Explicit3D/include/exDyna3D.h:75, virtual) is entered for step WARM+1 and
// stops when it is entered for step WARM+STEPS+1, so it covers STEPS whole
// iterations of the `while (time < stepEndTime)` loop (exDyna3D.cpp:86-100):
// the dt update with its min-edge pass and computeSate, output_results
// included (its VTK writer is stubbed, dataout_stub.cpp).
// Prints: elements, steps timed, seconds, OpenMP threads.
#include "exDyna3D.h"
#include "dataIn.h"
#include "spdlog/spdlog.h"

#include <omp.h>
#include <unistd.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>

namespace {

struct Stop {};

class Timed : public EnSC::exDyna3D {
public:
    long warm = 0, steps = 0, n = 0;
    std::chrono::steady_clock::time_point t0, t1;
    void computeSate() override {
        ++n;
        if (n == warm + 1) t0 = std::chrono::steady_clock::now();
        if (n == warm + steps + 1) {
            t1 = std::chrono::steady_clock::now();
            throw Stop();
        }
        EnSC::exDyna3D::computeSate();
    }
    size_t n_elems() const { return hexahedron_elements.size(); }
};

}  // namespace

int main(int argc, char** argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s WORKDIR DECK WARM STEPS [c_dr c_cr factor]\n", argv[0]);
        return 2;
    }
    const std::string work = argv[1], deck = argv[2];
    spdlog::set_level(spdlog::level::warn);
    if (chdir(work.c_str()) != 0) {
        std::perror("chdir");
        return 2;
    }
    {
        // project.txt is read from the working directory (exDyna3D.cpp:705-712)
        std::ofstream p("project.txt");
        p << "#deck\n" << deck << "\n";
        if (argc >= 8) p << "#\n#c_dr\n" << argv[5] << "\n#\n#c_cr\n" << argv[6] << "\n#\n#factor\n" << argv[7] << "\n";
    }
    Timed sim;
    sim.warm = std::atol(argv[3]);
    sim.steps = std::atol(argv[4]);
    EnSC::DataIn reader(sim);
    reader.read_inp(deck);
    try {
        sim.run();
    } catch (const Stop&) {
    }
    if (sim.n < sim.warm + sim.steps + 1) {
        std::fprintf(stderr, "the run ended after %ld steps\n", sim.n);
        return 1;
    }
    const double secs = std::chrono::duration<double>(sim.t1 - sim.t0).count();
    std::printf("elements %zu steps %ld seconds %.9g threads %d\n", sim.n_elems(), sim.steps, secs,
                omp_get_max_threads());
    return 0;
}
