// pandemic_sim.cpp — see pandemic_sim.h.  Citations are relative to
// PandemicThrust/PandemicThrust/ of the reference.
#include "pandemic_sim.h"

#include <sys/stat.h>

#include <cstdlib>
#include <cstring>

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); }

// Kernel classes of pt_sim_kernel_name, looked up by name so the host
// does not depend on their order.
struct Classes {
    int sweep, plan, compact, gen_wd, gen_we, up, scan, down, offsets, c_wd, c_we, apply, end_day, mark, select,
        subset;
    int count;
    Classes() {
        count = 0;
        while (count < 64 && pt_sim_kernel_name(count)[0]) ++count;
        auto find = [](const char* n) {
            for (int k = 0; k < 64; ++k) {
                const char* s = pt_sim_kernel_name(k);
                if (!s || !*s) break;
                if (strcmp(s, n) == 0) return k;
            }
            return -1;
        };
        sweep = find("sweep");
        plan = find("plan");
        compact = find("compact");
        gen_wd = find("gen_weekday");
        gen_we = find("gen_weekend");
        up = find("sort_upsweep");
        scan = find("sort_scan");
        down = find("sort_downsweep");
        offsets = find("offsets");
        c_wd = find("contacts_weekday");
        c_we = find("contacts_weekend");
        apply = find("apply");
        end_day = find("end_day");
        mark = find("mark");
        select = find("select");
        subset = find("subset_sort");
    }
};

bool file_exists(const std::string& path) {
    struct stat st;
    return stat(path.c_str(), &st) == 0;
}

}  // namespace

// ---------------- ResourceLog (resource_logging.cu) ----------------

void ResourceLog::setup(const SimOptions& o) {
    out_dir_ = o.output_dir.empty() ? o.input_dir : o.output_dir;
    device_ = o.device;
    if (pt_device_memory(device_, &initial_free_, &initial_total_) != PT_OK) {
        fprintf(stderr, "pandemicthrust: device memory: %s\n", pt_last_error());
        exit(1);
    }
    if (o.poll_memory) {
        const std::string path = out_dir_ + "/output_mem.csv";
        f_mem_ = fopen(path.c_str(), "w");
        if (!f_mem_) {
            perror("Error opening output_mem.csv");
            exit(1);
        }
        fprintf(f_mem_, "day,freeBytes,bytesUsed,totalBytes,megabytesUsed\n");
        fprintf(f_mem_, "INITIAL,%zu,0,%zu,0\n", (size_t)initial_free_, (size_t)initial_total_);
    }
    start_ = Clock::now();
}

void ResourceLog::sample(int day) {
    uint64_t fr = 0, tot = 0;
    if (pt_device_memory(device_, &fr, &tot) != PT_OK) return;
    const uint64_t used = initial_free_ > fr ? initial_free_ - fr : 0;
    if (used > max_used_) max_used_ = used;
    if (f_mem_) fprintf(f_mem_, "%d,%zu,%zu,%zu,%zu\n", day, (size_t)fr, (size_t)used, (size_t)tot, (size_t)(used >> 20));
}

void ResourceLog::done(float p_scale, float l_scale, int seed) {
    const float ms = (float)ms_since(start_);
    const float sec = ms / 1000;
    char dev[256] = "unknown";
    if (pt_device_name(device_, dev, sizeof dev) == PT_OK)
        for (char* c = dev; *c; ++c)
            if (*c == ',') *c = ' ';
    const std::string path = out_dir_ + "/output_resource_log.csv";
    const bool exists = file_exists(path);
    FILE* f = fopen(path.c_str(), exists ? "a" : "w");
    if (f) {
        if (!exists)
            fprintf(f, "sim_type,sim_device,people_sim_scale,location_sim_scale,seed,runtime_milliseconds,"
                       "runtime_seconds,bytes_used,megabytes_used\n");
        fprintf(f, "%s,%s,%f,%f,%d,%f,%f,%zu,%zu\n", "gfx950_hip", dev, p_scale, l_scale, seed, ms, sec,
                (size_t)max_used_, (size_t)(max_used_ >> 20));
        fclose(f);
    }
    if (f_mem_) fclose(f_mem_);
    f_mem_ = nullptr;
}

// ---------------- ProfileLog (profiler.cu) ----------------

bool ProfileLog::open(const std::string& path) {
    f_ = fopen(path.c_str(), "w");
    if (!f_) return false;
    fprintf(f_, "day,stack_depth,function_name,problem_size,inclusive_time_milliseconds,exclusive_time_milliseconds\n");
    return true;
}

void ProfileLog::row(int day, int depth, const char* name, int size, double inclusive_ms, double exclusive_ms) {
    if (!f_) return;
    fprintf(f_, "%d, %d, %s, %d, %f, %lf\n", day, depth, name, size, (double)(float)inclusive_ms, exclusive_ms);
}

void ProfileLog::close() {
    if (f_) fclose(f_);
    f_ = nullptr;
}

// ---------------- PandemicSim ----------------

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
    const auto t0 = Clock::now();
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
    load_params_ms_ = ms_since(t0);
}

// PandemicSim() (PandemicSim.cu:98-133): open the output streams, load
// inputs, scale and push the model (pt_sim_create).
PandemicSim::PandemicSim(const SimOptions& opts, ResourceLog* res)
    : opt_(opts), res_(res), dir_(opts.input_dir), out_dir_(opts.output_dir.empty() ? opts.input_dir : opts.output_dir) {
    const std::string stats = out("output_infected_stats.csv");
    f_stats_ = fopen(stats.c_str(), "w");
    if (!f_stats_) {
        perror("Error opening output_infected_stats.csv");
        exit(1);
    }
    fprintf(f_stats_, "day,pandemic_susceptible,pandemic_infectious,pandemic_symptomatic,pandemic_asymptomatic,"
                      "pandemic_recovered,seasonal_susceptible,seasonal_infectious,seasonal_symptomatic,"
                      "seasonal_asymptomatic,seasonal_recovered\n");
    if (opt_.profile && !prof_.open(out("profile_log.csv"))) {
        perror("Error opening profile_log.csv");
        exit(1);
    }
    if (opt_.validation) {
        val_.reset(new Validator);
        if (!val_->open(out_dir_)) {
            perror("Error opening validation files");
            exit(1);
        }
    }
    loadParameters();
    if (val_) val_->print("parameters loaded");
    const int32_t key[2] = {seeds_[0], seeds_[1]};
    if (pt_sim_create(&params_, key, opt_.device, &sim_) != PT_OK) fatal("create");
    if (val_ && pt_sim_enable_trace(sim_, 1) != PT_OK) fatal("trace");
    double ph[PT_SETUP_PHASES] = {0};
    pt_sim_setup_times(sim_, ph, PT_SETUP_PHASES);
    // Constructor-level functions sit at stack depth 0 (PandemicSim.cu:121-127).
    prof_.row(-1, 0, "setup_loadParameters", 1, load_params_ms_, load_params_ms_);
    prof_.row(-1, 0, "setup_scaleSimulation", 14, ph[PT_SETUP_SCALE], ph[PT_SETUP_SCALE]);
    prof_.row(-1, 0, "setup_pushDeviceData", 1, ph[PT_SETUP_PUSH], ph[PT_SETUP_PUSH]);
}

// ~PandemicSim: logging_closeOutputStreams (PandemicSim.cu:838-872).
PandemicSim::~PandemicSim() {
    if (!stopped_) {
        FILE* f = fopen(out("infected_max.txt").c_str(), "w");
        if (f) {
            fprintf(f, "max_proportion,day\n%f,%d\n", max_proportion_, max_proportion_day_);
            fclose(f);
        }
    }
    if (f_stats_) fclose(f_stats_);
    prof_.close();
    val_.reset();
    pt_sim_destroy(sim_);
}

// setupSim (PandemicSim.cu:146-208); with a checkpoint to resume from, the
// loaded state replaces setup_initialInfected.
void PandemicSim::setupSim() {
    const auto t0 = Clock::now();
    if (val_) val_->print("setting up households");
    if (pt_sim_setup(sim_, &number_people_) != PT_OK) fatal("setup");
    if (val_) val_->setup(sim_);  // debug_validatePeopleSetup (PandemicSim.cu:186-193)
    int64_t sc[8];
    if (pt_sim_get_scalars(sim_, sc) != PT_OK) fatal("setup");
    number_households_ = (int)sc[1];
    number_workplaces_ = (int)sc[2];
    printf("%d people, %d households, %d workplaces\n", (int)number_people_, number_households_, number_workplaces_);
    const auto t_seed = Clock::now();
    const int n = params_.init_pandemic + params_.init_seasonal;
    const char* seed_name = "setup_initialInfected";
    int seed_size = n;
    if (!opt_.resume_path.empty()) {
        seed_name = "setup_loadCheckpoint";
        FILE* f = fopen(opt_.resume_path.c_str(), "rb");
        if (!f) {
            perror("Error opening checkpoint");
            exit(1);
        }
        std::vector<uint8_t> buf((size_t)pt_sim_checkpoint_bytes(sim_));
        const size_t got = fread(buf.data(), 1, buf.size(), f);
        fclose(f);
        if (pt_sim_checkpoint_load(sim_, buf.data(), (int64_t)got) != PT_OK) fatal("resume");
        int64_t s2[8];
        pt_sim_get_scalars(sim_, s2);
        printf("resumed from %s at day %d\n", opt_.resume_path.c_str(), (int)s2[6]);
        seed_size = (int)s2[6];
    } else {
        std::vector<int32_t> ids((size_t)(n > 0 ? n : 1));
        if (pt_glibc_initial_infected(core_seed_, (int32_t)number_people_, n, ids.data()) != PT_OK)
            fatal("initial infected");
        if (pt_sim_seed_infections(sim_, ids.data(), params_.init_pandemic, params_.init_seasonal) != PT_OK)
            fatal("initial infected");
    }
    const double seed_ms = ms_since(t_seed);
    if (val_) val_->print("simulation setup complete");
    // must be done every simulation (PandemicSim.cu:195-196)
    res_->sample(0);
    double ph[PT_SETUP_PHASES] = {0};
    pt_sim_setup_times(sim_, ph, PT_SETUP_PHASES);
    const int N = (int)number_people_;
    prof_.row(-1, 1, "setup_calcPopulationSize_thrust", number_households_, ph[PT_SETUP_POPULATION_SIZE],
              ph[PT_SETUP_POPULATION_SIZE]);
    prof_.row(-1, 1, "setup_sizeGlobalArrays", N, ph[PT_SETUP_SIZE_ARRAYS], ph[PT_SETUP_SIZE_ARRAYS]);
    prof_.row(-1, 1, "setup_generateHouseholds", N, ph[PT_SETUP_HOUSEHOLDS], ph[PT_SETUP_HOUSEHOLDS]);
    prof_.row(-1, 1, "setup_assignWorkplaces", N, ph[PT_SETUP_WORKPLACES], ph[PT_SETUP_WORKPLACES]);
    prof_.row(0, 1, seed_name, seed_size, seed_ms, seed_ms);
    const double total = ms_since(t0);
    const double children = ph[PT_SETUP_POPULATION_SIZE] + ph[PT_SETUP_SIZE_ARRAYS] + ph[PT_SETUP_HOUSEHOLDS] +
                            ph[PT_SETUP_WORKPLACES] + seed_ms;
    prof_.row(-1, 0, "setupSim", N, total, total - children);
}

// daily_writeInfectedStats (PandemicSim.cu:2771-2818) for every row since
// setup; the console lines and the infected maximum (:901-931) follow the
// same rows.
void PandemicSim::writeStats(int rows, const int32_t* counts, const int32_t* infected, std::vector<double>* row_ms) {
    row_ms->assign((size_t)rows, 0.0);
    for (int d = 0; d < rows; ++d) {
        const auto t = Clock::now();
        const int32_t* c = &counts[(size_t)d * 16];
        const int psym = c[2] + c[3] + c[4], pasym = c[5] + c[6] + c[7];
        const int ssym = c[10] + c[11] + c[12], sasym = c[13] + c[14] + c[15];
        fprintf(f_stats_, "%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d\n", d, c[1], psym + pasym, psym, pasym, c[0], c[9],
                ssym + sasym, ssym, sasym, c[8]);
        (*row_ms)[(size_t)d] = ms_since(t);
        const int I = infected[d];
        if (I == 0) break;
        const float proportion = (float)I / (int)number_people_;
        if (proportion > max_proportion_) {
            max_proportion_ = proportion;
            max_proportion_day_ = d;
        }
    }
    fflush(f_stats_);
}

// Per-day profile rows in the reference's nesting (PandemicSim.cu:882-954,
// :1186-1420, :2207-2224, :2738-2825, :3025-3043), timed per kernel class.
void PandemicSim::profileDays(int first_day, int rows, const int32_t* infected, const std::vector<double>& write_ms,
                              double* children_ms) {
    static const Classes K;
    const int days = pt_sim_day_kernel_times(sim_, nullptr, 0);
    if (days <= 0 || K.count <= 0) return;
    std::vector<double> T((size_t)days * K.count);
    pt_sim_day_kernel_times(sim_, T.data(), days);
    auto t = [&](int d, int k) { return (k < 0 || d >= days) ? 0.0 : T[(size_t)d * K.count + k]; };
    const int N = (int)number_people_, W = number_workplaces_;
    for (int d = first_day; d < rows; ++d) {
        const int I = infected[d];
        const double w = write_ms[(size_t)d];
        const double sweep = t(d, K.sweep);
        prof_.row(d, 2, "daily_writeInfectedStats", 1, w, w);
        prof_.row(d, 1, "daily_countAndRecover", N, sweep + w, sweep);
        const double build = t(d, K.plan) + t(d, K.compact);
        prof_.row(d, 1, "daily_buildInfectedArray_global", I, build, build);
        *children_ms += sweep + w + build;
        if (I == 0) break;
        const double sort = t(d, K.mark) + t(d, K.select) + t(d, K.up) + t(d, K.scan) + t(d, K.down) + t(d, K.subset);
        const double search = t(d, K.offsets), end = t(d, K.end_day);
        double whole;
        if (d % 7 >= 5) {
            const double gen = t(d, K.gen_we), kern = t(d, K.c_we) + t(d, K.apply);
            prof_.row(d, 2, "weekend_assignErrands", N, gen, gen);
            prof_.row(d, 2, "doWeekend_wholeDay_setup_sort", N, sort, sort);
            prof_.row(d, 2, "doWeekend_wholeDay_setup_locationSearch", N, search, search);
            prof_.row(d, 2, "doWeekend_wholeDay_kernel", I, kern, kern);
            whole = gen + sort + search + kern + end;
            prof_.row(d, 1, "doWeekend_wholeDay", I, whole, end);
        } else {
            const double gen = t(d, K.gen_wd), kern = t(d, K.c_wd) + t(d, K.apply);
            prof_.row(d, 3, "weekday_generateAfterschoolAndErrandDestinations", N, gen, gen);
            prof_.row(d, 2, "doWeekday_wholeDay_generateSchedulesKernel", N, gen, 0.0);
            prof_.row(d, 2, "doWeekday_wholeDay_setup_sort", 2 * N, sort, sort);
            prof_.row(d, 2, "doWeekday_wholeDay_setup_locationSearch", 10 * W, search, search);
            prof_.row(d, 2, "doWeekday_wholeDay_kernel", I, kern, kern);
            whole = gen + sort + search + kern + end;
            prof_.row(d, 1, "doWeekday_wholeDay", I, whole, end);
        }
        *children_ms += whole;
    }
}

// runToCompletion (PandemicSim.cu:877-967) and final_countReproduction
// (:2227-2304).  Without per-day memory polling or a checkpoint the loop is
// enqueued on the GPU without per-day host synchronisation and the daily
// rows are written once it ends; otherwise it advances one day at a time.
void PandemicSim::runToCompletion() {
    const auto t0 = Clock::now();
    const int md = params_.max_days;
    int64_t sc[8];
    if (pt_sim_get_scalars(sim_, sc) != PT_OK) fatal("run");
    const int first_day = (int)sc[7];  // rows already produced (a resumed run)
    if (opt_.profile && pt_sim_enable_timing(sim_, -1) != PT_OK) fatal("timing");
    const bool stepwise = opt_.poll_memory || opt_.checkpoint_day >= 0 || val_;
    if (stepwise) {
        int32_t c[16], inf = 0;
        for (int d = first_day; d < md; ++d) {
            if (opt_.checkpoint_day >= 0 && d >= opt_.checkpoint_day) {
                stopped_ = true;
                break;
            }
            if (val_) val_->print("beginning day...");
            const int rc = pt_sim_step(sim_, c, &inf);
            if (rc < 0) fatal("step");
            if (inf == 0 || rc == 1) break;
            if (val_) val_->day(sim_, d, inf);
            if (opt_.poll_memory) res_->sample(d);
        }
    } else {
        int32_t days = 0;
        if (pt_sim_run(sim_, nullptr, nullptr, &days) != PT_OK) fatal("run");
    }
    std::vector<int32_t> counts((size_t)256 * 16), infected(256);
    const int rows = pt_sim_daily_rows(sim_, counts.data(), infected.data(), 256);
    if (rows < 0) fatal("rows");
    for (int d = first_day; d < rows && infected[d] > 0; ++d) printf("Day %3d:\tinfected: %5d\n", d + 1, infected[d]);
    std::vector<double> write_ms;
    writeStats(rows, counts.data(), infected.data(), &write_ms);
    double children = 0;
    if (opt_.profile) {
        profileDays(first_day, rows, infected.data(), write_ms, &children);
        pt_sim_enable_timing(sim_, 0);  // clears the day times read above
    }
    if (stopped_) {
        const int64_t bytes = pt_sim_checkpoint_bytes(sim_);
        std::vector<uint8_t> buf((size_t)bytes);
        if (pt_sim_checkpoint_save(sim_, buf.data(), bytes) != PT_OK) fatal("checkpoint");
        FILE* f = fopen(opt_.checkpoint_path.c_str(), "wb");
        if (!f || fwrite(buf.data(), 1, buf.size(), f) != buf.size()) {
            perror("Error writing checkpoint");
            exit(1);
        }
        fclose(f);
        printf("checkpoint written to %s after day %d\n", opt_.checkpoint_path.c_str(), rows);
        return;
    }
    // final_releaseMemory (:3958-3971): nothing is released here (the loop's
    // buffers are reused by the next run); the row keeps the memory sample.
    const auto t_rel = Clock::now();
    if (opt_.poll_memory) res_->sample(255);
    const double rel_ms = ms_since(t_rel);
    prof_.row(255, 1, "final_releaseMemory", 1, rel_ms, rel_ms);
    const auto t_final = Clock::now();
    std::vector<int32_t> cp((size_t)md + 1), cs((size_t)md + 1);
    if (pt_sim_generation_offsets(sim_, cp.data(), cs.data(), md + 1) != PT_OK) fatal("reproduction");
    const double search_ms = ms_since(t_final);
    const auto t_host = Clock::now();
    FILE* f = fopen(out("output_rn.csv").c_str(), "w");
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
    const double host_ms = ms_since(t_host);
    const double final_ms = ms_since(t_final);
    const int N = (int)number_people_;
    prof_.row(-1, 2, "final_countReproduction_genSearch", N, search_ms, search_ms);
    prof_.row(-1, 2, "final_countReproduction_hostCode", md, host_ms, host_ms);
    prof_.row(-1, 1, "final_countReproduction", N, final_ms, final_ms - search_ms - host_ms);
    const double total = ms_since(t0);
    prof_.row(-1, 0, "runToCompletion", N, total, total - children - rel_ms - final_ms);
}
