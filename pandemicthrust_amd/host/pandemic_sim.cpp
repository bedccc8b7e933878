// pandemic_sim.cpp — see pandemic_sim.h.  Citations are relative to
// PandemicThrust/PandemicThrust/ of the reference.
#include "pandemic_sim.h"

#include <cstdlib>
#include <cstring>

void PandemicSim::fatal(const char* what) {
    fprintf(stderr, "pandemicthrust: %s: %s\n", what, pt_last_error());
    exit(1);
}

// setup_loadSeed (PandemicSim.cu:2843-2862): missing input -> perror, exit 1.
void PandemicSim::loadSeed() {
    const std::string path = dir_ + "/seed.txt";
    FILE* f = fopen(path.c_str(), "r");
    if (!f) {
        perror("Error opening seed file");
        exit(1);
    }
    if (fscanf(f, "%d", &core_seed_) != 1) {
        fprintf(stderr, "Error reading seed file: expected an integer\n");
        exit(1);
    }
    fclose(f);
    if (pt_glibc_seeds(core_seed_, seeds_) != PT_OK) fatal("seeds");
}

// setup_loadParameters (PandemicSim.cu:268-323): positional parse of the
// second line of constants.csv after the label column.
void PandemicSim::loadParameters() {
    loadSeed();
    printf("seeds:\t%d\t%d\t%d\t%d\n", seeds_[0], seeds_[1], seeds_[2], seeds_[3]);
    const std::string path = dir_ + "/constants.csv";
    FILE* f = fopen(path.c_str(), "r");
    if (!f) {
        perror("Error opening constants file");
        exit(1);
    }
    char line[512];
    if (!fgets(line, sizeof line, f)) {
        fprintf(stderr, "Error reading constants file: empty\n");
        exit(1);
    }
    int got = 0;
    if (fscanf(f, "%*[^,]%*c") < 0) got = -100;
    got += fscanf(f, "%d%*c", &params_.max_days) == 1;
    got += fscanf(f, "%f%*c", &params_.r_pandemic) == 1;
    got += fscanf(f, "%f%*c", &params_.r_seasonal) == 1;
    got += fscanf(f, "%d%*c", &params_.init_pandemic) == 1;
    got += fscanf(f, "%d%*c", &params_.init_seasonal) == 1;
    got += fscanf(f, "%f%*c", &params_.people_scaling) == 1;
    got += fscanf(f, "%f%*c", &params_.location_scaling) == 1;
    got += fscanf(f, "%f%*c", &params_.percent_symptomatic) == 1;
    got += fscanf(f, "%f", &params_.asymp_factor) == 1;
    fclose(f);
    if (got != 9) {
        fprintf(stderr, "Error reading constants file: malformed data row\n");
        exit(1);
    }
    printf("max days: %d\nr_p: %f\nr_s: %f\ninitial_pandemic: %d\ninitial_seasonal: %d\nnumber_households: %d\n",
           params_.max_days, params_.r_pandemic, params_.r_seasonal, params_.init_pandemic, params_.init_seasonal,
           1000000);
}

// PandemicSim() (PandemicSim.cu:98-133): open the stats file, load inputs,
// scale and push the model (pt_sim_create).
PandemicSim::PandemicSim(const std::string& dir, int device) : dir_(dir), device_(device) {
    const std::string stats = dir_ + "/output_infected_stats.csv";
    f_stats_ = fopen(stats.c_str(), "w");
    if (!f_stats_) {
        perror("Error opening output_infected_stats.csv");
        exit(1);
    }
    fprintf(f_stats_, "day,pandemic_susceptible,pandemic_infectious,pandemic_symptomatic,pandemic_asymptomatic,"
                      "pandemic_recovered,seasonal_susceptible,seasonal_infectious,seasonal_symptomatic,"
                      "seasonal_asymptomatic,seasonal_recovered\n");
    loadParameters();
    const int32_t key[2] = {seeds_[0], seeds_[1]};
    if (pt_sim_create(&params_, key, device_, &sim_) != PT_OK) fatal("create");
}

// ~PandemicSim: logging_closeOutputStreams (PandemicSim.cu:838-872).
PandemicSim::~PandemicSim() {
    const std::string path = dir_ + "/infected_max.txt";
    FILE* f = fopen(path.c_str(), "w");
    if (f) {
        fprintf(f, "max_proportion,day\n%f,%d\n", max_proportion_, max_proportion_day_);
        fclose(f);
    }
    if (f_stats_) fclose(f_stats_);
    pt_sim_destroy(sim_);
}

// setupSim (PandemicSim.cu:146-208).
void PandemicSim::setupSim() {
    if (pt_sim_setup(sim_, &number_people_) != PT_OK) fatal("setup");
    int64_t sc[8];
    if (pt_sim_get_scalars(sim_, sc) != PT_OK) fatal("setup");
    number_households_ = (int)sc[1];
    number_workplaces_ = (int)sc[2];
    printf("%d people, %d households, %d workplaces\n", (int)number_people_, number_households_, number_workplaces_);
    const int n = params_.init_pandemic + params_.init_seasonal;
    std::vector<int32_t> ids((size_t)(n > 0 ? n : 1));
    if (pt_glibc_initial_infected(core_seed_, (int32_t)number_people_, n, ids.data()) != PT_OK)
        fatal("initial infected");
    if (pt_sim_seed_infections(sim_, ids.data(), params_.init_pandemic, params_.init_seasonal) != PT_OK)
        fatal("initial infected");
}

// runToCompletion (PandemicSim.cu:877-967) and final_countReproduction
// (:2227-2304).  The loop is enqueued on the GPU without per-day host
// synchronisation; the daily rows are written once it ends.
void PandemicSim::runToCompletion() {
    const int md = params_.max_days;
    std::vector<int32_t> counts((size_t)(md > 0 ? md : 1) * 16), infected((size_t)(md > 0 ? md : 1));
    int32_t days = 0;
    if (pt_sim_run(sim_, counts.data(), infected.data(), &days) != PT_OK) fatal("run");
    loop_ms_ = pt_sim_last_run_ms(sim_);
    for (int d = 0; d < days; ++d) {
        const int32_t* c = &counts[(size_t)d * 16];
        const int psym = c[2] + c[3] + c[4], pasym = c[5] + c[6] + c[7];
        const int ssym = c[10] + c[11] + c[12], sasym = c[13] + c[14] + c[15];
        fprintf(f_stats_, "%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d\n", d, c[1], psym + pasym, psym, pasym, c[0], c[9],
                ssym + sasym, ssym, sasym, c[8]);
        const int I = infected[d];
        if (I == 0) break;
        printf("Day %3d:\tinfected: %5d\n", d + 1, I);
        const float proportion = (float)I / (int)number_people_;
        if (proportion > max_proportion_) {
            max_proportion_ = proportion;
            max_proportion_day_ = d;
        }
    }
    fflush(f_stats_);
    std::vector<int32_t> cp((size_t)md + 1), cs((size_t)md + 1);
    if (pt_sim_generation_offsets(sim_, cp.data(), cs.data(), md + 1) != PT_OK) fatal("reproduction");
    const std::string path = dir_ + "/output_rn.csv";
    FILE* f = fopen(path.c_str(), "w");
    if (!f) {
        perror("Error opening output_rn.csv");
        exit(1);
    }
    fprintf(f, "gen,gen_size_p,rn_p,gen_size_s,rn_s\n");
    for (int gen = 0; gen < md - 1; ++gen) {
        const int size_p = cp[gen + 1] - cp[gen], next_p = cp[gen + 2] - cp[gen + 1];
        const float rn_p = (float)next_p / size_p;
        const int size_s = cs[gen + 1] - cs[gen], next_s = cs[gen + 2] - cs[gen + 1];
        const float rn_s = (float)next_s / size_s;
        fprintf(f, "%d,%d,%f,%d,%f\n", gen, size_p, rn_p, size_s, rn_s);
    }
    fclose(f);
}
