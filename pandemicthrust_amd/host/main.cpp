// main.cpp — the pandemicthrust executable (reference: main.cu:12-31).
// Reads constants.csv and seed.txt from the working directory (or from the
// directory given as the only argument) and writes the reference's output
// files there.
#include "pandemic_sim.h"

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    PandemicSim sim(dir);
    sim.setupSim();
    sim.runToCompletion();
    return 0;
}
