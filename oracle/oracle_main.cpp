/*
 * CPU ORACLE driver — TEST INFRASTRUCTURE ONLY.
 *
 *   pt_oracle                 run like the reference binary: reads
 *                             constants.csv and seed.txt from the working
 *                             directory, writes the output files there.
 *   pt_oracle --time DAYS     timing mode for the CPU baseline: runs setup
 *                             and at most DAYS days, prints one JSON line.
 */
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "pt_oracle.h"

/* setup_loadParameters / setup_loadSeed parsing (PandemicSim.cu:287-310, :2845-2854) */
static int load_inputs(or_params* p, int* seed) {
    FILE* fs = fopen("seed.txt", "r");
    if (!fs) {
        perror("Error opening seed file");
        return 1;
    }
    if (fscanf(fs, "%d", seed) != 1) {
        fclose(fs);
        fprintf(stderr, "seed.txt: expected an integer\n");
        return 1;
    }
    fclose(fs);
    FILE* fc = fopen("constants.csv", "r");
    if (!fc) {
        perror("Error opening constants file");
        return 1;
    }
    char line[512];
    if (!fgets(line, sizeof line, fc)) {
        fclose(fc);
        return 1;
    }
    int ok = 0;
    ok += fscanf(fc, "%*[^,]%*c") >= 0;
    ok += fscanf(fc, "%d%*c", &p->max_days) == 1;
    ok += fscanf(fc, "%f%*c", &p->r_pandemic) == 1;
    ok += fscanf(fc, "%f%*c", &p->r_seasonal) == 1;
    ok += fscanf(fc, "%d%*c", &p->init_pandemic) == 1;
    ok += fscanf(fc, "%d%*c", &p->init_seasonal) == 1;
    ok += fscanf(fc, "%f%*c", &p->people_scaling) == 1;
    ok += fscanf(fc, "%f%*c", &p->location_scaling) == 1;
    ok += fscanf(fc, "%f%*c", &p->percent_symptomatic) == 1;
    ok += fscanf(fc, "%f", &p->asymp_factor) == 1;
    fclose(fc);
    if (ok != 10) {
        fprintf(stderr, "constants.csv: malformed data row\n");
        return 1;
    }
    return 0;
}

int main(int argc, char** argv) {
    int time_days = -1;
    if (argc >= 3 && std::strcmp(argv[1], "--time") == 0) time_days = std::atoi(argv[2]);
    or_params p{};
    int seed = 0;
    if (load_inputs(&p, &seed)) return 1;
    or_sim* s = or_create(&p, seed);
    if (!s) {
        fprintf(stderr, "oracle: %s\n", or_last_error());
        return 1;
    }
    auto t0 = std::chrono::steady_clock::now();
    if (or_setup(s) != 0) {
        fprintf(stderr, "oracle: %s\n", or_last_error());
        return 1;
    }
    auto t1 = std::chrono::steady_clock::now();
    const int cap = time_days >= 0 ? time_days : p.max_days;
    std::vector<int32_t> counts((size_t)(cap > 0 ? cap : 1) * 16), inf((size_t)(cap > 0 ? cap : 1));
    const int days = or_run(s, counts.data(), inf.data(), cap);
    auto t2 = std::chrono::steady_clock::now();
    const double setup_s = std::chrono::duration<double>(t1 - t0).count();
    const double loop_s = std::chrono::duration<double>(t2 - t1).count();
    const long long N = or_get_scalar(s, 0);
    if (time_days >= 0) {
        printf("{\"n_people\": %lld, \"days\": %d, \"setup_s\": %.6f, \"loop_s\": %.6f, \"person_days_per_s\": %.3f}\n",
               N, days, setup_s, loop_s, loop_s > 0 ? (double)N * days / loop_s : 0.0);
    } else {
        for (int d = 0; d < days; ++d)
            if (inf[d] > 0) printf("Day %3d:\tinfected: %5d\n", d + 1, inf[d]);
        if (or_write_outputs(s, ".", counts.data(), inf.data(), days) != 0) {
            fprintf(stderr, "oracle: cannot write outputs\n");
            return 1;
        }
    }
    or_destroy(s);
    return 0;
}
