// pandemic_sim.h — host driver with the reference's PandemicSim flow
// (PandemicThrust/PandemicThrust/PandemicSim.h:68-281, main.cu:12-31):
// the constructor reads seed.txt and constants.csv from the input
// directory, setupSim() builds the population on the GPU, runToCompletion()
// runs the daily loop and writes output_infected_stats.csv and
// output_rn.csv; the destructor writes infected_max.txt.  Alongside, as the
// reference does with its default switches (PandemicSim.h:23-45), it writes
// profile_log.csv (profiler.cu), appends output_resource_log.csv
// (resource_logging.cu:83-122) and, on request, output_mem.csv
// (resource_logging.cu:31-81), and in validation mode debug.txt and the
// debug CSVs (validation.h).  All device work goes through the C ABI of
// include/pandemic/pt_abi.h.
#pragma once

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pandemic/pt_abi.h"
#include "validation.h"

struct SimOptions {
    std::string input_dir = ".";
    std::string output_dir;         // empty: the input directory
    int device = 0;
    bool profile = true;            // SIM_PROFILING (PandemicSim.h:39)
    bool poll_memory = false;       // POLL_MEMORY_USAGE (PandemicSim.h:33)
    std::string checkpoint_path;    // write a checkpoint here ...
    int checkpoint_day = -1;        // ... once this many days have run, then stop
    std::string resume_path;        // continue from this checkpoint instead of seeding
    bool validation = false;        // SIM_VALIDATION (PandemicSim.h:50): debug.txt and the debug CSVs
};

// Memory samples and the run clock of resource_logging.cu.
class ResourceLog {
public:
    void setup(const SimOptions& o);                        // logging_pollMemUsage_doSetup
    void sample(int day);                                   // logging_pollMemoryUsage_takeSample
    void done(float p_scale, float l_scale, int seed);      // logging_pollMemoryUsage_done

private:
    std::string out_dir_;
    int device_ = 0;
    FILE* f_mem_ = nullptr;
    uint64_t initial_free_ = 0, initial_total_ = 0, max_used_ = 0;
    std::chrono::steady_clock::time_point start_;
};

// Rows of profile_log.csv: "day, stack_depth, function_name, problem_size,
// inclusive ms, exclusive ms" (profiler.cu:13, :72-76).
class ProfileLog {
public:
    bool open(const std::string& path);
    void row(int day, int depth, const char* name, int size, double inclusive_ms, double exclusive_ms);
    void close();

private:
    FILE* f_ = nullptr;
};

class PandemicSim {
public:
    explicit PandemicSim(const SimOptions& opts, ResourceLog* res);
    ~PandemicSim();

    void setupSim();
    void runToCompletion();

    int64_t number_people() const { return number_people_; }
    float people_scaling() const { return params_.people_scaling; }
    float location_scaling() const { return params_.location_scaling; }
    int core_seed() const { return core_seed_; }
    bool stopped_at_checkpoint() const { return stopped_; }

private:
    void fatal(const char* what);
    void loadSeed();
    void loadParameters();
    void writeStats(int rows, const int32_t* counts, const int32_t* infected, std::vector<double>* row_ms);
    void profileDays(int first_day, int rows, const int32_t* infected, const std::vector<double>& write_ms,
                     double* children_ms);
    std::string out(const char* name) const { return out_dir_ + "/" + name; }

    SimOptions opt_;
    ResourceLog* res_ = nullptr;
    std::string dir_, out_dir_;
    int core_seed_ = 0;
    int32_t seeds_[4] = {0, 0, 0, 0};
    pt_params params_{};
    pt_sim* sim_ = nullptr;
    int64_t number_people_ = 0;
    int number_households_ = 0, number_workplaces_ = 0;
    FILE* f_stats_ = nullptr;
    ProfileLog prof_;
    double load_params_ms_ = 0;
    float max_proportion_ = 0.f;
    int max_proportion_day_ = 0;
    bool stopped_ = false;
    std::unique_ptr<Validator> val_;
};
