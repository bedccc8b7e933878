// pandemic_sim.h — host driver with the reference's PandemicSim flow
// (PandemicThrust/PandemicThrust/PandemicSim.h:68-281, main.cu:12-31):
// the constructor reads seed.txt and constants.csv from the working
// directory, setupSim() builds the population on the GPU, runToCompletion()
// runs the daily loop and writes output_infected_stats.csv and
// output_rn.csv; the destructor writes infected_max.txt.  All device work
// goes through the C ABI of include/pandemic/pt_abi.h.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/pandemic/pt_abi.h"

class PandemicSim {
public:
    // Reads the inputs from dir (the reference reads the working directory).
    explicit PandemicSim(const std::string& dir = ".", int device = 0);
    ~PandemicSim();

    void setupSim();
    void runToCompletion();

    int64_t number_people() const { return number_people_; }
    double loop_ms() const { return loop_ms_; }

private:
    void fatal(const char* what);
    void loadSeed();
    void loadParameters();

    std::string dir_;
    int device_ = 0;
    int core_seed_ = 0;
    int32_t seeds_[4] = {0, 0, 0, 0};
    pt_params params_{};
    pt_sim* sim_ = nullptr;
    int64_t number_people_ = 0;
    int number_households_ = 0, number_workplaces_ = 0;
    FILE* f_stats_ = nullptr;
    float max_proportion_ = 0.f;
    int max_proportion_day_ = 0;
    double loop_ms_ = 0;
};
