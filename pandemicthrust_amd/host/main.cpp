// main.cpp — the pandemicthrust executable (reference: main.cu:12-31).
//
//   pandemicthrust [options] [input_dir]
//
// Reads constants.csv and seed.txt from input_dir (default: the working
// directory) and writes the reference's output files there (or into
// --output-dir).  Options:
//   --output-dir DIR        where the output files go
//   --device K              HIP device ordinal
//   --no-profile            do not write profile_log.csv
//   --poll-memory           write output_mem.csv (a sample every day)
//   --checkpoint FILE --checkpoint-day D
//                           stop once D days have run and save the loop
//                           state to FILE (output_rn.csv and
//                           infected_max.txt are then not written)
//   --resume FILE           continue a run from FILE instead of seeding
//   --validation            validation mode: debug.txt, debug_infected.csv,
//                           debug_contacts.csv and debug_actions.csv
#include <cstdlib>
#include <cstring>

#include "pandemic_sim.h"

static void usage() {
    fprintf(stderr,
            "usage: pandemicthrust [--output-dir DIR] [--device K] [--no-profile] [--poll-memory]\n"
            "                      [--checkpoint FILE --checkpoint-day D] [--resume FILE]\n"
            "                      [--validation] [input_dir]\n");
    exit(2);
}

int main(int argc, char** argv) {
    SimOptions o;
    bool have_dir = false;
    for (int i = 1; i < argc; ++i) {
        const char* a = argv[i];
        auto value = [&]() -> const char* {
            if (i + 1 >= argc) usage();
            return argv[++i];
        };
        if (!strcmp(a, "--output-dir")) o.output_dir = value();
        else if (!strcmp(a, "--device")) o.device = atoi(value());
        else if (!strcmp(a, "--no-profile")) o.profile = false;
        else if (!strcmp(a, "--poll-memory")) o.poll_memory = true;
        else if (!strcmp(a, "--checkpoint")) o.checkpoint_path = value();
        else if (!strcmp(a, "--checkpoint-day")) o.checkpoint_day = atoi(value());
        else if (!strcmp(a, "--resume")) o.resume_path = value();
        else if (!strcmp(a, "--validation")) o.validation = true;
        else if (a[0] == '-' || have_dir) usage();
        else {
            o.input_dir = a;
            have_dir = true;
        }
    }
    if ((o.checkpoint_day >= 0) != !o.checkpoint_path.empty()) usage();
    ResourceLog res;
    res.setup(o);  // logging_pollMemUsage_doSetup
    {
        PandemicSim sim(o, &res);
        sim.setupSim();
        sim.runToCompletion();
        if (!sim.stopped_at_checkpoint()) res.done(sim.people_scaling(), sim.location_scaling(), sim.core_seed());
    }
    return 0;
}
