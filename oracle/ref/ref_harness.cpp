/*
 * Golden-vector harness — TEST INFRASTRUCTURE ONLY.
 *
 * Built by oracle/ref/Makefile against the reference's own sources where they
 * lie under /root/reference: PandemicThrust/PandemicThrust/host_functions.cpp
 * (the reference's CPU twins of its schedule generators) and the vendored
 * Random123 threefry.h.  The globals those files expect from PandemicSim.cu
 * are defined here; their values come from the reference's
 * setup_loadParameters() through the generated ref_tables.h.  Output: CSV
 * fixtures for tests/golden/.
 *
 * usage: ref_harness <output directory>
 */
#include "stdafx.h"
#include "simParameters.h"
#include "host_functions.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

/* Globals that host_functions.h declares extern (host_functions.h:28-37). */
float WORKPLACE_TYPE_ASSIGNMENT_PDF_HOST[NUM_BUSINESS_TYPES];
float WORKPLACE_TYPE_WEEKDAY_ERRAND_PDF_HOST[NUM_BUSINESS_TYPES];
float WORKPLACE_TYPE_WEEKEND_ERRAND_PDF_HOST[NUM_BUSINESS_TYPES];
int WORKPLACE_TYPE_OFFSET_HOST[NUM_BUSINESS_TYPES];
int WORKPLACE_TYPE_COUNT_HOST[NUM_BUSINESS_TYPES];
float CHILD_AGE_CDF_HOST[CHILD_DATA_ROWS];
int CHILD_AGE_SCHOOLTYPE_LOOKUP_HOST[CHILD_DATA_ROWS];
simRandOffsetsStruct_t host_randOffsetsStruct[1];
simSizeConstantsStruct_t host_simSizeStruct[1];
SEED_T SEED_HOST[SEED_LENGTH];

/* Other tables assigned in setup_loadParameters (dumped for comparison). */
int WORKPLACE_TYPE_MAX_CONTACTS_HOST[NUM_BUSINESS_TYPES];
float HOUSEHOLD_TYPE_CDF_HOST[HH_TABLE_ROWS];
int HOUSEHOLD_TYPE_ADULT_COUNT_HOST[HH_TABLE_ROWS];
int HOUSEHOLD_TYPE_CHILD_COUNT_HOST[HH_TABLE_ROWS];
int WEEKDAY_ERRAND_CONTACT_ASSIGNMENT_HOST[4][2];
int WEEKEND_ERRAND_CONTACT_ASSIGNMENTS_HOST[6][2];
float VIRAL_SHEDDING_PROFILES_HOST[NUM_SHEDDING_PROFILES][CULMINATION_PERIOD];

#include "ref_tables.h"

static int base_count[NUM_BUSINESS_TYPES];

/* Restatement of setup_scaleSimulation's workplace part (PandemicSim.cu:1135-1158),
 * needed because that code lives in the CUDA translation unit. */
static int scale_workplaces(float lscale) {
    int sum = 0;
    for (int t = 0; t < NUM_BUSINESS_TYPES; ++t) {
        int orig = (int)floor((double)base_count[t] + 0.5);
        int nc = (int)floor((double)(lscale * orig) + 0.5);
        if (nc == 0 && orig > 0) nc = 1;
        WORKPLACE_TYPE_COUNT_HOST[t] = nc;
        sum += nc;
    }
    int acc = 0;
    for (int t = 0; t < NUM_BUSINESS_TYPES; ++t) {
        WORKPLACE_TYPE_OFFSET_HOST[t] = acc;
        acc += WORKPLACE_TYPE_COUNT_HOST[t];
    }
    return sum;
}

static uint32_t lcg_state = 12345u;
static uint32_t lcg() {
    lcg_state = lcg_state * 1664525u + 1013904223u;
    return lcg_state;
}

static FILE* open_out(const std::string& dir, const char* name) {
    std::string p = dir + "/" + name;
    FILE* f = fopen(p.c_str(), "w");
    if (!f) {
        perror(p.c_str());
        exit(1);
    }
    return f;
}

static void write_threefry(const std::string& dir) {
    FILE* f = open_out(dir, "threefry_ref.csv");
    fprintf(f, "kind,c0,c1,k0,k1,o0,o1\n");
    const uint64_t in64[][4] = {
        {0ull, 0ull, 0ull, 0ull},
        {~0ull, ~0ull, ~0ull, ~0ull},
        {0x243f6a8885a308d3ull, 0x13198a2e03707344ull, 0xa4093822299f31d0ull, 0x082efa98ec4e6c89ull},
    };
    for (auto& v : in64) {
        threefry2x64_ctr_t c = {{v[0], v[1]}};
        threefry2x64_key_t k = {{v[2], v[3]}};
        threefry2x64_ctr_t r = threefry2x64(c, k);
        fprintf(f, "2x64,%llu,%llu,%llu,%llu,%llu,%llu\n", (unsigned long long)v[0], (unsigned long long)v[1],
                (unsigned long long)v[2], (unsigned long long)v[3], (unsigned long long)r.v[0],
                (unsigned long long)r.v[1]);
    }
    const uint32_t in32[][4] = {
        {0u, 0u, 0u, 0u}, {~0u, ~0u, ~0u, ~0u}, {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u}};
    for (auto& v : in32) {
        threefry2x32_ctr_t c = {{v[0], v[1]}};
        threefry2x32_key_t k = {{v[2], v[3]}};
        threefry2x32_ctr_t r = threefry2x32(c, k);
        fprintf(f, "2x32,%u,%u,%u,%u,%u,%u\n", v[0], v[1], v[2], v[3], r.v[0], r.v[1]);
    }
    /* The reference's calling convention: ctr {off,off}, key {SEED[0],SEED[1]}. */
    const int seeds[] = {1, 2, 12345};
    for (int sd : seeds) {
        srand(sd);
        int s0 = rand(), s1 = rand();
        const uint64_t offs[] = {0ull, 1ull, 5310000ull, 4294967295ull, 4294967296ull, 123456789012ull};
        for (uint64_t off : offs) {
            threefry2x64_key_t k = {{(uint64_t)(long)s0, (uint64_t)(long)s1}};
            threefry2x64_ctr_t c = {{off, off}};
            threefry2x64_ctr_t r = threefry2x64(c, k);
            fprintf(f, "2x64,%llu,%llu,%llu,%llu,%llu,%llu\n", (unsigned long long)off, (unsigned long long)off,
                    (unsigned long long)k.v[0], (unsigned long long)k.v[1], (unsigned long long)r.v[0],
                    (unsigned long long)r.v[1]);
            threefry2x32_key_t k32 = {{(uint32_t)s0, (uint32_t)s1}};
            threefry2x32_ctr_t c32 = {{(uint32_t)(off + 1), (uint32_t)(off + 1)}};
            threefry2x32_ctr_t r32 = threefry2x32(c32, k32);
            fprintf(f, "2x32,%u,%u,%u,%u,%u,%u\n", c32.v[0], c32.v[1], k32.v[0], k32.v[1], r32.v[0], r32.v[1]);
        }
    }
    for (int i = 0; i < 64; ++i) {
        uint64_t c0 = ((uint64_t)lcg() << 32) | lcg(), c1 = ((uint64_t)lcg() << 32) | lcg();
        uint64_t k0 = ((uint64_t)lcg() << 32) | lcg(), k1 = ((uint64_t)lcg() << 32) | lcg();
        threefry2x64_ctr_t c = {{c0, c1}};
        threefry2x64_key_t k = {{k0, k1}};
        threefry2x64_ctr_t r = threefry2x64(c, k);
        fprintf(f, "2x64,%llu,%llu,%llu,%llu,%llu,%llu\n", (unsigned long long)c0, (unsigned long long)c1,
                (unsigned long long)k0, (unsigned long long)k1, (unsigned long long)r.v[0],
                (unsigned long long)r.v[1]);
        uint32_t a = lcg(), b = lcg(), x = lcg(), y = lcg();
        threefry2x32_ctr_t c32 = {{a, b}};
        threefry2x32_key_t k32 = {{x, y}};
        threefry2x32_ctr_t r32 = threefry2x32(c32, k32);
        fprintf(f, "2x32,%u,%u,%u,%u,%u,%u\n", a, b, x, y, r32.v[0], r32.v[1]);
    }
    fclose(f);
}

static void write_tables(const std::string& dir) {
    FILE* f = open_out(dir, "ref_tables.csv");
    fprintf(f, "name,i,j,value\n");
    for (int i = 0; i < NUM_BUSINESS_TYPES; ++i) {
        fprintf(f, "wp_pdf,%d,0,%.9g\n", i, WORKPLACE_TYPE_ASSIGNMENT_PDF_HOST[i]);
        fprintf(f, "wd_pdf,%d,0,%.9g\n", i, WORKPLACE_TYPE_WEEKDAY_ERRAND_PDF_HOST[i]);
        fprintf(f, "we_pdf,%d,0,%.9g\n", i, WORKPLACE_TYPE_WEEKEND_ERRAND_PDF_HOST[i]);
        fprintf(f, "wp_count,%d,0,%d\n", i, base_count[i]);
        fprintf(f, "wp_max,%d,0,%d\n", i, WORKPLACE_TYPE_MAX_CONTACTS_HOST[i]);
    }
    for (int i = 0; i < HH_TABLE_ROWS; ++i) {
        fprintf(f, "hh_cdf,%d,0,%.9g\n", i, HOUSEHOLD_TYPE_CDF_HOST[i]);
        fprintf(f, "hh_adults,%d,0,%d\n", i, HOUSEHOLD_TYPE_ADULT_COUNT_HOST[i]);
        fprintf(f, "hh_children,%d,0,%d\n", i, HOUSEHOLD_TYPE_CHILD_COUNT_HOST[i]);
    }
    for (int i = 0; i < CHILD_DATA_ROWS; ++i) {
        fprintf(f, "child_cdf,%d,0,%.9g\n", i, CHILD_AGE_CDF_HOST[i]);
        fprintf(f, "child_school,%d,0,%d\n", i, CHILD_AGE_SCHOOLTYPE_LOOKUP_HOST[i]);
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 2; ++j) fprintf(f, "wd_assign,%d,%d,%d\n", i, j, WEEKDAY_ERRAND_CONTACT_ASSIGNMENT_HOST[i][j]);
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 2; ++j) fprintf(f, "we_assign,%d,%d,%d\n", i, j, WEEKEND_ERRAND_CONTACT_ASSIGNMENTS_HOST[i][j]);
    for (int i = 0; i < NUM_SHEDDING_PROFILES; ++i)
        for (int j = 0; j < CULMINATION_PERIOD; ++j) fprintf(f, "shed,%d,%d,%.9g\n", i, j, VIRAL_SHEDDING_PROFILES_HOST[i][j]);
    fclose(f);
}

/* Rand values that probe the CDF/PDF walk boundaries plus a random sweep. */
static int probe_values(uint32_t* out, int cap) {
    int n = 0;
    const uint32_t fixed[] = {0u, 1u, 2u, 127u, 128u, 0x7FFFFFFFu, 0x80000000u, 0xFFFFFF00u,
                              0xFFFFFF7Fu, 0xFFFFFF80u, 0xFFFFFFFEu, 0xFFFFFFFFu};
    for (uint32_t v : fixed) out[n++] = v;
    const float* tabs[] = {WORKPLACE_TYPE_ASSIGNMENT_PDF_HOST, WORKPLACE_TYPE_WEEKDAY_ERRAND_PDF_HOST,
                           WORKPLACE_TYPE_WEEKEND_ERRAND_PDF_HOST};
    for (const float* t : tabs) {
        double cum = 0;
        for (int i = 0; i < NUM_BUSINESS_TYPES; ++i) {
            cum += t[i];
            double b = cum * 4294967296.0;
            for (int d = -300; d <= 300; d += 60) {
                double v = b + d;
                if (v >= 0 && v < 4294967296.0 && n < cap) out[n++] = (uint32_t)v;
            }
        }
    }
    for (int i = 0; i < CHILD_DATA_ROWS; ++i) {
        double b = (double)CHILD_AGE_CDF_HOST[i] * 4294967296.0;
        for (int d = -300; d <= 300; d += 60) {
            double v = b + d;
            if (v >= 0 && v < 4294967296.0 && n < cap) out[n++] = (uint32_t)v;
        }
    }
    while (n < cap) out[n++] = lcg();
    return n;
}

static void write_schedules(const std::string& dir) {
    struct Cfg { const char* name; float lscale; int core_seed; uint64_t wp_off; uint64_t er_off; };
    const Cfg cfgs[] = {
        {"small_100k", 0.0425f, 1, 10625ull, 45630ull},
        {"default", 1.0f, 1, 250000ull, 838000ull},
        {"full_50M", 21.24f, 1, 5310000ull, 17809740ull},
        {"tiny", 0.001f, 7, 0ull, 1ull},
        {"wrap", 3.0f, 12345, 4294967290ull, 4294967200ull},
    };
    FILE* fw = open_out(dir, "host_workplace.csv");
    FILE* fd = open_out(dir, "host_weekday.csv");
    FILE* fe = open_out(dir, "host_weekend.csv");
    FILE* ff = open_out(dir, "host_fish.csv");
    FILE* fc = open_out(dir, "ref_scaling.csv");
    fprintf(fw, "cfg,lscale,core_seed,workplace_off,idx,age,workplace\n");
    fprintf(fd, "cfg,lscale,core_seed,errand_off,idx,age,dest1,dest2,profile\n");
    fprintf(fe, "cfg,lscale,core_seed,errand_off,idx,d0,d1,d2\n");
    fprintf(ff, "cfg,lscale,r,workplace,child_age,school,wd_adult,wd_child,weekend\n");
    fprintf(fc, "cfg,lscale,type,count,offset,W\n");
    static uint32_t probes[768];
    const int np = probe_values(probes, 768);
    for (const Cfg& c : cfgs) {
        const int W = scale_workplaces(c.lscale);
        for (int t = 0; t < NUM_BUSINESS_TYPES; ++t)
            fprintf(fc, "%s,%.9g,%d,%d,%d,%d\n", c.name, c.lscale, t, WORKPLACE_TYPE_COUNT_HOST[t],
                    WORKPLACE_TYPE_OFFSET_HOST[t], W);
        srand(c.core_seed);
        for (int i = 0; i < SEED_LENGTH; ++i) SEED_HOST[i] = rand();
        host_randOffsetsStruct->workplace_randOffset = c.wp_off;
        host_randOffsetsStruct->errand_randOffset = c.er_off;
        host_simSizeStruct->number_workplaces = W;
        const int idxs[] = {0, 1, 2, 3, 4, 5, 6, 7, 99, 100, 101, 102, 103, 12345, 99999, 100000,
                            1048575, 2354000, 49998959, 123456789};
        for (int idx : idxs) {
            for (int adult = 0; adult < 2; ++adult) {
                const age_t age = adult ? AGE_ADULT : (age_t)(idx % 5);
                fprintf(fw, "%s,%.9g,%d,%llu,%d,%d,%d\n", c.name, c.lscale, c.core_seed,
                        (unsigned long long)c.wp_off, idx, age, host_recalcWorkplace(idx, age));
                locId_t d1, d2;
                const int prof = host_recalc_weekdayErrandDests_assignProfile(idx, age, &d1, &d2);
                fprintf(fd, "%s,%.9g,%d,%llu,%d,%d,%d,%d,%d\n", c.name, c.lscale, c.core_seed,
                        (unsigned long long)c.er_off, idx, age, d1, d2, prof);
            }
            locId_t e[3];
            host_recalc_weekendErrandDests(idx, e);
            fprintf(fe, "%s,%.9g,%d,%llu,%d,%d,%d,%d\n", c.name, c.lscale, c.core_seed,
                    (unsigned long long)c.er_off, idx, e[0], e[1], e[2]);
        }
        for (int i = 0; i < np; ++i) {
            const uint32_t r = probes[i];
            age_t child_age = 0;
            locId_t school = 0;
            host_setup_fishSchoolAndAge(r, &child_age, &school);
            fprintf(ff, "%s,%.9g,%u,%d,%d,%d,%d,%d,%d\n", c.name, c.lscale, r, host_setup_fishWorkplace(r),
                    child_age, school, host_fishAfterschoolOrErrandDestination_weekday(r, AGE_ADULT),
                    host_fishAfterschoolOrErrandDestination_weekday(r, 0), host_fishWeekendErrandDestination(r));
        }
    }
    fclose(fw);
    fclose(fd);
    fclose(fe);
    fclose(ff);
    fclose(fc);
}

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s <output dir>\n", argv[0]);
        return 2;
    }
    ref_load_tables();
    memcpy(base_count, WORKPLACE_TYPE_COUNT_HOST, sizeof(base_count));
    const std::string dir = argv[1];
    write_threefry(dir);
    write_tables(dir);
    write_schedules(dir);
    return 0;
}
