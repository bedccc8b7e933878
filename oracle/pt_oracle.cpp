/*
 * CPU ORACLE — TEST INFRASTRUCTURE ONLY (see pt_oracle.h).
 *
 * A sequential, single-threaded restatement of the reference daily loop.
 * Every function cites the reference lines it follows; paths are relative to
 * PandemicThrust/PandemicThrust/ unless noted.  Floating point follows the
 * reference's single-precision operation order exactly; the build uses
 * -ffp-contract=off and SSE arithmetic so no operation is fused or widened.
 * Float->int conversions use GPU semantics (truncate, saturate, NaN->0).
 *
 * How each part is pinned (tests/test_oracle_golden.py and the parity tests):
 *   Threefry             published Random123 KATs + vectors of the vendored header
 *   model tables         every literal of the reference's setup_loadParameters()
 *   scaling, workplace, weekday/weekend errands, fishing walks
 *                        vectors of the reference's own host_functions.cpp
 *   household-type walk  numpy inverse CDF over the reference's literals (ties,
 *                        neighbours of every boundary, random draws)
 *   tables, counts, R    closed forms (stable bucket order, lower_bound, totals)
 *   contacts, infection  per-contact binomial rates against shed * factor /
 *                        max(kval_sum, 1) and the R calibration, profile
 *                        special cases (pct = 1, pct = 0, y == pct), the
 *                        two-success profile property of the xor word, the
 *                        min-(position, slot) winner rule, R = 0, the
 *                        reference's output file contract
 *                        (tests/test_oracle_infection_pins.py)
 *
 * OR_PLANT (compile-time, tests only): 0 = the restatement.  Values 1..8
 * build a copy with one deliberate, plausible mistake in the infection
 * arithmetic so that the tests can show each pin fails on it:
 *   1 profile threshold y >= pct instead of y > pct
 *   2 shedding value read one day late
 *   3 profile word u[i] ^ u[(i + 2) % 16] instead of u[i] ^ u[(i + 1) % 16]
 *   4 contact weights normalised by kval_sum + 1
 *   5 strain factor R * (1 - a) * pct instead of R / ((1 - a) * pct)
 *   6 victim generation = infector generation (no + 1)
 *   7 an infector's slots processed last to first
 *   8 profile word u[i] (no xor)
 */
#ifndef OR_PLANT
#define OR_PLANT 0
#endif
#include "pt_oracle.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

/* ---- model constants (PandemicSim.cu:329-481, simParameters.h) ---- */
constexpr int NUM_TYPES = 14;
constexpr int HH_ROWS = 9;
constexpr int CHILD_ROWS = 5;
constexpr int AGE_ADULT = 5;
constexpr int CULMINATION = 10;
constexpr int FIRST_ERRAND_ROW = 9;
constexpr int TYPE_AFTERSCHOOL = 8;
constexpr int TYPE_PRESCHOOL = 3;
constexpr int WD_HOURS = 2, WE_HOURS = 10;
constexpr int WD_ERRANDS = 2, WE_ERRANDS = 3;
constexpr int C_WD = 8, C_WE = 5;
enum { CT_NONE = 0, CT_WORKPLACE, CT_SCHOOL, CT_ERRAND, CT_AFTERSCHOOL, CT_HOME };
constexpr uint8_t ST_RECOVERED = 0, ST_SUSCEPTIBLE = 1, ST_INFECTED = 2;
constexpr uint8_t DAY_NONE = 255, GEN_NONE = 255;

const float HH_CDF[HH_ROWS] = {0.279f, 0.319f, 0.628f, 0.671f, 0.8f, 0.812f, 0.939f, 0.944f, 1.0f};
const int HH_ADULTS[HH_ROWS] = {1, 1, 2, 1, 2, 1, 2, 1, 2};
const int HH_CHILDREN[HH_ROWS] = {0, 1, 0, 2, 1, 3, 2, 4, 3};
const float CHILD_CDF[CHILD_ROWS] = {0.24f, 0.47f, 0.72f, 0.85f, 1.0f};
const float WP_PDF[NUM_TYPES] = {0.06586f, 0.05802f, 0.30227f, 0.0048f, 0.00997f, 0.203f, 0.09736f,
                                 0.10598f, 0.00681f, 0.02599f, 0.f, 0.08749f, 0.03181f, 0.00064f};
const int WP_COUNT_BASE[NUM_TYPES] = {1000, 7000, 2400, 300, 100, 200, 100, 100, 300, 500, 0, 300, 400, 100};
const int WP_MAX[NUM_TYPES] = {3, 3, 3, 2, 2, 3, 3, 2, 2, 2, 0, 2, 2, 2};
const float WD_PDF[NUM_TYPES] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0.61919f, 0, 0.27812f, 0.06601f, 0.03668f};
const float WE_PDF[NUM_TYPES] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0.51493f, 0, 0.25586f, 0.1162f, 0.113f};
/* PandemicSim.cu:445-456: contacts per errand hour for each weekday profile. */
const int WD_ASSIGN[4][2] = {{2, 0}, {0, 2}, {1, 1}, {2 /* WP_MAX[afterschool] */, 0}};
/* PandemicSim.cu:460-481: which errand each of the two weekend contacts uses. */
const int WE_ASSIGN[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
const float SHED_BASE[6][10] = {
    {0.007339835f, 0.332600216f, 0.501192066f, 0.142183447f, 0.015675154f, 0.000967407f, 4.055E-05f,
     1.29105E-06f, 3.34836E-08f, 7.41011E-10f},
    {0.002533572f, 0.348252834f, 0.498210218f, 0.130145145f, 0.018421298f, 0.002158374f, 0.000245489f,
     2.88922E-05f, 3.61113E-06f, 4.83901E-07f},
    {0.05927385f, 0.314171688f, 0.411588802f, 0.187010054f, 0.026934715f, 0.001013098f, 7.78449E-06f,
     9.29441E-09f, 1.29796E-12f, 0},
    {0.04687299f, 0.248505983f, 0.30307952f, 0.211008627f, 0.11087006f, 0.049241932f, 0.019562658f,
     0.007179076f, 0.002482875f, 0.000820094f},
    {0.028667712f, 0.283445338f, 0.319240133f, 0.190123057f, 0.093989959f, 0.044155659f, 0.020682822f,
     0.009841839f, 0.00479234f, 0.002393665f},
    {0.087866042f, 0.223005225f, 0.258992749f, 0.208637267f, 0.127489076f, 0.061148649f, 0.023406737f,
     0.007216643f, 0.001802145f, 0.00036581f}};
const int KVAL[6] = {0, 1, 1, 1, 1, 1};

/* ---- Threefry (Salmon et al., SC'11; Skein rotation constants) ---- */
inline uint64_t rotl64(uint64_t x, unsigned n) { return (x << n) | (x >> (64 - n)); }
inline uint32_t rotl32(uint32_t x, unsigned n) { return (x << n) | (x >> (32 - n)); }

void tf2x64(uint64_t c0, uint64_t c1, uint64_t k0, uint64_t k1, uint64_t out[2]) {
    static const unsigned R[8] = {16, 42, 12, 31, 16, 32, 24, 21};
    const uint64_t ks[3] = {k0, k1, 0x1BD11BDAA9FC1A22ull ^ k0 ^ k1};
    uint64_t x0 = c0 + ks[0], x1 = c1 + ks[1];
    for (int r = 0; r < 20; ++r) {
        x0 += x1;
        x1 = rotl64(x1, R[r & 7]);
        x1 ^= x0;
        if ((r & 3) == 3) {
            const int s = (r + 1) >> 2;
            x0 += ks[s % 3];
            x1 += ks[(s + 1) % 3] + (uint64_t)s;
        }
    }
    out[0] = x0;
    out[1] = x1;
}

void tf2x32(uint32_t c0, uint32_t c1, uint32_t k0, uint32_t k1, uint32_t out[2]) {
    static const unsigned R[8] = {13, 15, 26, 6, 17, 29, 16, 24};
    const uint32_t ks[3] = {k0, k1, 0x1BD11BDAu ^ k0 ^ k1};
    uint32_t x0 = c0 + ks[0], x1 = c1 + ks[1];
    for (int r = 0; r < 20; ++r) {
        x0 += x1;
        x1 = rotl32(x1, R[r & 7]);
        x1 ^= x0;
        if ((r & 3) == 3) {
            const int s = (r + 1) >> 2;
            x0 += ks[s % 3];
            x1 += ks[(s + 1) % 3] + (uint32_t)s;
        }
    }
    out[0] = x0;
    out[1] = x1;
}

/* The reference reads Threefry output through a union of u32 words
 * (e.g. PandemicSim.cu:2332-2341): word 2i is the low half of 64-bit
 * output i, word 2i+1 the high half (little endian). */
struct Words4 { uint32_t w[4]; };
inline Words4 draw4(uint64_t ctr, uint64_t k0, uint64_t k1) {
    uint64_t o[2];
    tf2x64(ctr, ctr, k0, k1, o);
    Words4 r;
    r.w[0] = (uint32_t)o[0];
    r.w[1] = (uint32_t)(o[0] >> 32);
    r.w[2] = (uint32_t)o[1];
    r.w[3] = (uint32_t)(o[1] >> 32);
    return r;
}

/* (float)u / UNSIGNED_MAX, where UNSIGNED_MAX converts to 2^32 in float
 * (e.g. PandemicSim.cu:2363). */
inline float unit_float(uint32_t u) {
    volatile float f = (float)u;
    volatile float d = (float)(unsigned)0xFFFFFFFFu;
    return f / d;
}
/* f32 -> i32 with GPU (v_cvt_i32_f32) semantics. */
inline int32_t cvt_i32(float f) {
    if (f != f) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f < -2147483648.0f) return (int32_t)0x80000000;
    return (int32_t)f;
}
/* f32 -> u32 with GPU (v_cvt_u32_f32) semantics. */
inline uint32_t cvt_u32(float f) {
    if (f != f) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    if (f <= 0.0f) return 0u;
    return (uint32_t)f;
}
inline float fmul(float a, float b) { volatile float r = a * b; return r; }
inline float fdiv(float a, float b) { volatile float r = a / b; return r; }
inline float fsub(float a, float b) { volatile float r = a - b; return r; }

int round_half_up(double d) { return (int)std::floor(d + 0.5); } /* PandemicSim.cu:1662 */

}  // namespace

struct or_sim {
    or_params p{};
    int32_t core_seed = 0;
    int32_t seed[4]{};
    uint64_t k0 = 0, k1 = 0;
    /* scaled tables */
    int H = 0, W = 0, N = 0;
    int count[NUM_TYPES]{}, offset[NUM_TYPES]{};
    float shed[6][10]{};
    float factor[2]{};
    float pct = 0.f;
    /* counters */
    uint64_t rand_offset = 0, workplace_off = 0, errand_off = 0;
    int current_day = 0;
    bool done = false;
    int days_simulated = 0;
    bool setup_done = false;
    /* state */
    std::vector<uint8_t> hh_types, age, st_p, st_s, day_p, day_s, gen_p, gen_s;
    std::vector<int32_t> hh_offsets, household, wp_offsets, wp_people, wp_keys, infected;
    std::vector<int32_t> errand_offsets, errand_people;
    int errand_hours_today = 0;
    /* interventions */
    or_interventions iv{{0, 0}, {CULMINATION, CULMINATION}, 0, 0, 0.f, 0.f, {0.f, 0.f}};
    uint64_t ik0 = 0, ik1 = 0;
    /* trace */
    bool trace_on = false;
    std::vector<int32_t> tr_infector, tr_victim;
    std::vector<uint8_t> tr_type, tr_action, tr_draw, tr_prof;
};

namespace {

/* ---- setup_scaleSimulation (PandemicSim.cu:1127-1162) and
 *      setup_calculateInfectionData (PandemicSim.cu:2828-2840) ---- */
void scale_and_calibrate(or_sim* s) {
    s->H = round_half_up((double)s->p.people_scaling * (double)1000000);
    int sum = 0;
    for (int t = 0; t < NUM_TYPES; ++t) {
        const int orig = round_half_up((double)WP_COUNT_BASE[t]);
        volatile float prod = s->p.location_scaling * (float)orig; /* float * int */
        int nc = round_half_up((double)prod);
        if (nc == 0 && orig > 0) nc = 1;
        s->count[t] = nc;
        sum += nc;
    }
    s->W = sum;
    int acc = 0;
    for (int t = 0; t < NUM_TYPES; ++t) {
        s->offset[t] = acc;
        acc += s->count[t];
    }
    std::memcpy(s->shed, SHED_BASE, sizeof(SHED_BASE));
    for (int i = 3; i < 6; ++i)
        for (int j = 0; j < 10; ++j) s->shed[i][j] = fmul(s->shed[i][j], s->p.asymp_factor);
    const float r[2] = {s->p.r_pandemic, s->p.r_seasonal};
    for (int i = 0; i < 2; ++i) {
        volatile float a = 1.0f - s->p.asymp_factor;
        volatile float d = a * s->p.percent_symptomatic;
#if OR_PLANT == 5
        s->factor[i] = fmul(r[i], d);
#else
        s->factor[i] = fdiv(r[i], d);
#endif
    }
    s->pct = s->p.percent_symptomatic;
}

/* device_setup_fishHouseholdType (PandemicSim.cu:2360-2371) */
uint8_t fish_household(uint32_t r) {
    const float y = unit_float(r);
    int row = 0;
    while (y > HH_CDF[row] && row < HH_ROWS - 1) row++;
    return (uint8_t)row;
}

/* device_setup_fishWorkplace (PandemicSim.cu:2374-2399) */
int32_t fish_workplace(const or_sim* s, uint32_t r) {
    float y = unit_float(r);
    int row = 0;
    while (WP_PDF[row] < y && row < NUM_TYPES - 1) {
        y = fsub(y, WP_PDF[row]);
        row++;
    }
    const float frac = fdiv(y, WP_PDF[row]);
    const int type_count = s->count[row];
    int bn = cvt_i32(fmul(frac, (float)type_count));
    if (bn >= type_count) bn = type_count - 1;
    return bn + s->offset[row];
}

/* device_setup_fishSchoolAndAge (PandemicSim.cu:2402-2442) */
void fish_school_and_age(const or_sim* s, uint32_t r, uint8_t* age, int32_t* school) {
    const float y = unit_float(r);
    int row = 0;
    while (row < CHILD_ROWS - 1 && y > CHILD_CDF[row]) row++;
    const int wp_type = row + TYPE_PRESCHOOL;
    float frac;
    if (row == 0) {
        frac = fdiv(y, CHILD_CDF[row]);
    } else {
        const float pdf_here = fsub(CHILD_CDF[row], CHILD_CDF[row - 1]);
        const float y_here = fsub(y, CHILD_CDF[row - 1]);
        frac = fdiv(y_here, pdf_here);
    }
    const int type_count = s->count[wp_type];
    int bn = cvt_i32(fmul(frac, (float)type_count));
    if (bn >= type_count) bn = type_count - 1;
    *school = bn + s->offset[wp_type];
    *age = (uint8_t)row;
}

/* device_setup_assignWorkplaceOrSchool (PandemicSim.cu:2445-2457) */
void assign_workplace_or_school(const or_sim* s, uint32_t r, uint8_t* age, int32_t* wp) {
    if (*age == AGE_ADULT)
        *wp = fish_workplace(s, r);
    else
        fish_school_and_age(s, r, age, wp);
}

/* device_fishAfterschoolOrErrandDestination_weekday (PandemicSim.cu:2909-2947) */
int32_t fish_weekday_dest(const or_sim* s, uint32_t r, uint8_t age) {
    int bt = TYPE_AFTERSCHOOL;
    float frac = unit_float(r);
    if (age == AGE_ADULT) {
        bt = FIRST_ERRAND_ROW;
        float row_pdf = WD_PDF[bt];
        while (frac > row_pdf && bt < NUM_TYPES - 1) {
            frac = fsub(frac, row_pdf);
            bt++;
            row_pdf = WD_PDF[bt];
        }
        frac = fdiv(frac, row_pdf);
    }
    const int type_count = s->count[bt];
    const int type_offset = s->offset[bt];
    uint32_t bn = cvt_u32(fmul(frac, (float)type_count));
    if (bn >= (uint32_t)type_count) bn = (uint32_t)(type_count - 1);
    bn += (uint32_t)type_offset;
    return (int32_t)bn;
}

/* device_assignAfterschoolOrErrandDests_weekday (PandemicSim.cu:2950-2969) */
void weekday_dests(const or_sim* s, uint32_t r1, uint32_t r2, uint8_t age, int32_t* d1, int32_t* d2) {
    int a = fish_weekday_dest(s, r1, age);
    int b = a;
    if (age == AGE_ADULT) b = fish_weekday_dest(s, r2, age);
    b += s->W;
    *d1 = a;
    *d2 = b;
}

/* device_fishWeekendErrandDestination (PandemicSim.cu:1017-1044) */
int32_t fish_weekend_dest(const or_sim* s, uint32_t r) {
    float y = unit_float(r);
    int row = FIRST_ERRAND_ROW;
    while (y > WE_PDF[row] && row < NUM_TYPES - 1) {
        y = fsub(y, WE_PDF[row]);
        row++;
    }
    const float frac = fdiv(y, WE_PDF[row]);
    const int type_count = s->count[row];
    int bn = cvt_i32(fmul(frac, (float)type_count));
    if (bn >= type_count) bn = type_count - 1;
    return bn + s->offset[row];
}

/* device_generateWeekendErrands (PandemicSim.cu:980-1014) */
void weekend_errands(const or_sim* s, uint64_t off, int32_t out[3]) {
    const Words4 a = draw4(off, s->k0, s->k1);
    const Words4 b = draw4(off + 1, s->k0, s->k1);
    const uint32_t u[8] = {a.w[0], a.w[1], a.w[2], a.w[3], b.w[0], b.w[1], b.w[2], b.w[3]};
    const int h1 = (int)(u[0] % WE_HOURS);
    int h2 = (int)(u[1] % WE_HOURS);
    if (h2 == h1) h2 = (h2 + 1) % WE_HOURS;
    int h3 = (int)(u[2] % WE_HOURS);
    while (h3 == h1 || h3 == h2) h3 = (h3 + 1) % WE_HOURS;
    out[0] = fish_weekend_dest(s, u[3]) + h1 * s->W;
    out[1] = fish_weekend_dest(s, u[4]) + h2 * s->W;
    out[2] = fish_weekend_dest(s, u[5]) + h3 * s->W;
}

/* device_recalcWorkplace (PandemicSim.cu:3872-3890) */
int32_t recalc_workplace(const or_sim* s, int32_t idx, uint8_t age) {
    const uint64_t off = s->workplace_off + (uint64_t)(idx / 4);
    const Words4 u = draw4(off, s->k0, s->k1);
    int32_t wp;
    uint8_t a = age;
    assign_workplace_or_school(s, u.w[idx % 4], &a, &wp);
    return wp;
}

/* device_recalc_weekdayErrandDests_assignProfile (PandemicSim.cu:3667-3705) */
int recalc_weekday(const or_sim* s, int32_t idx, uint8_t age, int32_t* d1, int32_t* d2) {
    const uint64_t off = s->errand_off + (uint64_t)(idx / 2);
    const Words4 u = draw4(off, s->k0, s->k1);
    const int slot = 2 * (idx % 2);
    weekday_dests(s, u.w[slot], u.w[slot + 1], age, d1, d2);
    if (age != AGE_ADULT) return 3;
    const uint32_t x = u.w[slot] ^ u.w[slot + 1];
    return (int)(x % 3u);
}

/* device_recalc_weekendErrandDests (PandemicSim.cu:3742-3746) */
void recalc_weekend(const or_sim* s, int32_t idx, int32_t out[3]) {
    const uint64_t off = s->errand_off + (uint64_t)(int64_t)(2 * idx);
    weekend_errands(s, off, out);
}

/* device_getWorkplaceMaxContacts (PandemicSim.cu:3492-3511) */
int workplace_max_contacts(const or_sim* s, int32_t loc) {
    const int id = loc % s->W;
    int t = 0, off, cnt, mx;
    do {
        off = s->offset[t];
        cnt = s->count[t];
        mx = WP_MAX[t];
        t++;
    } while (id >= off + cnt);
    return mx;
}

/* device_selectRandomPersonFromLocation (PandemicSim.cu:2028-2071) with
 * device_getVictimAtIndex (:2016-2025).  Returns the contact's kval. */
int select_person(int32_t infector, int32_t loc_off, int32_t loc_count, uint32_t r, int type,
                  const int32_t* people, int32_t* victim_out, uint8_t* type_out) {
    int32_t victim = loc_count;
    int ct = CT_NONE;
    if (loc_count > 1) {
        int vo = (int)(r % (uint32_t)loc_count);
        victim = (type == CT_HOME) ? loc_off + vo : people[loc_off + vo];
        if (victim == infector) {
            vo = (vo + 1) % loc_count;
            victim = (type == CT_HOME) ? loc_off + vo : people[loc_off + vo];
        }
        ct = type;
    }
    *victim_out = victim;
    *type_out = (uint8_t)ct;
    return KVAL[ct];
}

/* device_makeContacts_weekday (PandemicSim.cu:1669-1829) */
int make_contacts_weekday(const or_sim* s, int32_t idx, uint8_t age, uint64_t M, int32_t* vic, uint8_t* typ) {
    const Words4 a = draw4(M, s->k0, s->k1);
    const Words4 b = draw4(M + 1, s->k0, s->k1);
    const uint32_t u[8] = {a.w[0], a.w[1], a.w[2], a.w[3], b.w[0], b.w[1], b.w[2], b.w[3]};
    int kval = 0;
    {
        const int hh = s->household[idx];
        const int off = s->hh_offsets[hh];
        const int cnt = s->hh_offsets[hh + 1] - off;
        for (int c = 0; c < 3; ++c) kval += select_person(idx, off, cnt, u[c], CT_HOME, nullptr, vic + c, typ + c);
    }
    int made = 3;
    {
        const int wp = recalc_workplace(s, idx, age);
        const int off = s->wp_offsets[wp];
        const int cnt = s->wp_offsets[wp + 1] - off;
        int desired = workplace_max_contacts(s, wp);
        int local = made;
        made += desired;
        const int kt = (age != AGE_ADULT) ? CT_SCHOOL : CT_WORKPLACE;
        while (desired > 0) {
            kval += select_person(idx, off, cnt, u[local], kt, s->wp_people.data(), vic + local, typ + local);
            desired--;
            local++;
        }
    }
    {
        const int kt = (age != AGE_ADULT) ? CT_AFTERSCHOOL : CT_ERRAND;
        int32_t d[2];
        const int prof = recalc_weekday(s, idx, age, d, d + 1);
        for (int hour = 0; hour < WD_HOURS; ++hour) {
            int desired = WD_ASSIGN[prof][hour];
            const int off = s->errand_offsets[d[hour]];
            const int cnt = s->errand_offsets[d[hour] + 1] - off;
            while (desired > 0 && made < C_WD) {
                kval += select_person(idx, off, cnt, u[made], kt, s->errand_people.data(), vic + made, typ + made);
                desired--;
                made++;
            }
        }
        while (made < C_WD) {
            vic[made] = -1;
            typ[made] = CT_NONE;
            made++;
        }
    }
    return kval;
}

/* device_makeContacts_weekend (PandemicSim.cu:1833-1950) */
int make_contacts_weekend(const or_sim* s, int32_t idx, uint64_t M, int32_t* vic, uint8_t* typ) {
    const Words4 u = draw4(M, s->k0, s->k1);
    int kval = 0;
    {
        const int hh = s->household[idx];
        const int off = s->hh_offsets[hh];
        const int cnt = s->hh_offsets[hh + 1] - off;
        for (int c = 0; c < 3; ++c) kval += select_person(idx, off, cnt, u.w[c], CT_HOME, nullptr, vic + c, typ + c);
    }
    const int prof = (int)(u.w[3] % 6u);
    uint32_t v[2];
    const uint32_t c32 = (uint32_t)(M + 1);
    tf2x32(c32, c32, (uint32_t)s->seed[0], (uint32_t)s->seed[1], v);
    int32_t d[3];
    recalc_weekend(s, idx, d);
    for (int k = 0; k < 2; ++k) {
        const int loc = d[WE_ASSIGN[prof][k]];
        const int off = s->errand_offsets[loc];
        const int cnt = s->errand_offsets[loc + 1] - off;
        kval += select_person(idx, off, cnt, v[k], CT_ERRAND, s->errand_people.data(), vic + 3 + k, typ + 3 + k);
    }
    return kval;
}

/* device_calculateInfectionProbability (PandemicSim.cu:2307-2326) */
float infection_probability(const or_sim* s, int profile, int dd, int strain, int kval_sum) {
    if (kval_sum == 0) kval_sum = 1;
#if OR_PLANT == 4
    kval_sum += 1;
#endif
#if OR_PLANT == 2
    const int flat = profile * CULMINATION + dd + 1;
    const float alpha = (&s->shed[0][0])[flat < 6 * CULMINATION ? flat : 6 * CULMINATION - 1];
#else
    const float alpha = s->shed[profile][dd];
#endif
    const float avg = fmul(alpha, s->factor[strain]);
    return fdiv(avg, (float)kval_sum);
}

/* device_getInfectionProfile (PandemicSim.cu:3974-3988) */
uint8_t infection_profile(const or_sim* s, uint32_t r) {
    uint32_t prof = r % 3u;
    prof += ST_INFECTED;
    const float y = unit_float(r);
#if OR_PLANT == 1
    if (y >= s->pct) prof += 3;
#else
    if (y > s->pct) prof += 3;
#endif
    return (uint8_t)prof;
}

/* device_processContacts (PandemicSim.cu:3106-3233) with the infection
 * applied immediately in sequence (device_doInfectionActionImmediately,
 * :3070-3103, and device_setInfectionStatus, :3048-3067). */
void process_contacts(or_sim* s, int32_t idx, int kval_sum, const int32_t* vic, const uint8_t* typ, int C,
                      uint64_t M2, int64_t trace_base) {
    const int d = s->current_day;
    const uint8_t sp = s->st_p[idx], ss = s->st_s[idx];
    float prob_p = -1.f, prob_s = -1.f;
    uint8_t gp = GEN_NONE, gs = GEN_NONE;
    if (sp >= ST_INFECTED) {
        const int pd = d - (int)s->day_p[idx] - s->iv.latent_days[0];
        if (pd >= 0 && pd < s->iv.infectious_days[0]) {
            prob_p = infection_probability(s, sp - ST_INFECTED, pd, 0, kval_sum);
            gp = (uint8_t)(s->gen_p[idx] + 1);
        }
    }
    if (ss >= ST_INFECTED) {
        const int pd = d - (int)s->day_s[idx] - s->iv.latent_days[1];
        if (pd >= 0 && pd < s->iv.infectious_days[1]) {
            prob_s = infection_probability(s, ss - ST_INFECTED, pd, 1, kval_sum);
            gs = (uint8_t)(s->gen_s[idx] + 1);
        }
    }
    uint32_t u[16];
    for (int k = 0; k < 4; ++k) {
        const Words4 w = draw4(M2 + (uint64_t)k, s->k0, s->k1);
        for (int j = 0; j < 4; ++j) u[4 * k + j] = w.w[j];
    }
#if OR_PLANT == 6
    if (gp != GEN_NONE) gp = (uint8_t)(gp - 1);
    if (gs != GEN_NONE) gs = (uint8_t)(gs - 1);
#endif
    const uint8_t day_to_set = (uint8_t)(d + 1);
    for (int k = 0; k < C; ++k) {
#if OR_PLANT == 7
        const int c = C - 1 - k;
#else
        const int c = k;
#endif
        int used = 2 * c; /* the reference's running rand_vals_used */
        const int32_t victim = vic[c];
        const int kv = KVAL[typ[c]];
        const float yp = unit_float(u[used]);
        const bool inf_p = yp < fmul(prob_p, (float)kv);
#if OR_PLANT == 3
        const uint8_t pp = infection_profile(s, u[used] ^ u[(used + 2) % 16]);
#elif OR_PLANT == 8
        const uint8_t pp = infection_profile(s, u[used]);
#else
        const uint8_t pp = infection_profile(s, u[used] ^ u[(used + 1) % 16]);
#endif
        used++;
        const float ys = unit_float(u[used]);
        const bool inf_s = ys < fmul(prob_s, (float)kv);
#if OR_PLANT == 3
        const uint8_t ps = infection_profile(s, u[used] ^ u[(used + 2) % 16]);
#elif OR_PLANT == 8
        const uint8_t ps = infection_profile(s, u[used]);
#else
        const uint8_t ps = infection_profile(s, u[used] ^ u[(used + 1) % 16]);
#endif
        used++;
        int action = 0;
        if (inf_p && s->st_p[victim] == ST_SUSCEPTIBLE) {
            s->st_p[victim] = pp;
            s->day_p[victim] = day_to_set;
            s->gen_p[victim] = gp;
            action += 2;
        }
        if (inf_s && s->st_s[victim] == ST_SUSCEPTIBLE) {
            s->st_s[victim] = ps;
            s->day_s[victim] = day_to_set;
            s->gen_s[victim] = gs;
            action += 1;
        }
        if (trace_base >= 0) {
            s->tr_infector[trace_base + c] = idx;
            s->tr_victim[trace_base + c] = victim;
            s->tr_type[trace_base + c] = typ[c];
            s->tr_action[trace_base + c] = (uint8_t)action;
            s->tr_draw[trace_base + c] = (uint8_t)((inf_p ? 2 : 0) | (inf_s ? 1 : 0));
            s->tr_prof[trace_base + c] = (uint8_t)(pp | (ps << 4));
        }
    }
}

/* Stable bucketing of (key, id) pairs whose ids appear in the order given:
 * equal to the stable radix sort plus thrust::lower_bound of the reference
 * (PandemicSim.cu:1212-1248, :1316-1354, :3835-3862). */
void bucket_stable(const std::vector<int32_t>& keys, const std::vector<int32_t>& ids, int nkeys,
                   std::vector<int32_t>& offsets, std::vector<int32_t>& people) {
    const size_t n = keys.size();
    offsets.assign((size_t)nkeys + 1, 0);
    for (size_t j = 0; j < n; ++j) offsets[(size_t)keys[j] + 1]++;
    for (int k = 0; k < nkeys; ++k) offsets[(size_t)k + 1] += offsets[k];
    std::vector<int32_t> cursor(offsets.begin(), offsets.end() - 1);
    people.resize(n);
    for (size_t j = 0; j < n; ++j) people[(size_t)cursor[keys[j]]++] = ids[j];
}

/* daily_countAndRecover / kernel_countInfectedStatusAndRecover
 * (PandemicSim.cu:2738-2768, :2097-2192). */
void count_and_recover(or_sim* s, int32_t counts[16]) {
    for (int i = 0; i < 16; ++i) counts[i] = 0;
    const int rec_p = s->current_day - (s->iv.latent_days[0] + s->iv.infectious_days[0]);
    const int rec_s = s->current_day - (s->iv.latent_days[1] + s->iv.infectious_days[1]);
    for (int i = 0; i < s->N; ++i) {
        int sp = s->st_p[i];
        if (rec_p >= 0 && sp >= ST_INFECTED && (int)s->day_p[i] <= rec_p) {
            sp = ST_RECOVERED;
            s->st_p[i] = ST_RECOVERED;
        }
        int ss = s->st_s[i];
        if (rec_s >= 0 && ss >= ST_INFECTED && (int)s->day_s[i] <= rec_s) {
            ss = ST_RECOVERED;
            s->st_s[i] = ST_RECOVERED;
        }
        counts[sp]++;
        counts[8 + ss]++;
    }
}

/* daily_buildInfectedArray_global (PandemicSim.cu:2207-2224): stable. */
void build_infected(or_sim* s) {
    s->infected.clear();
    for (int i = 0; i < s->N; ++i)
        if (s->st_p[i] >= ST_INFECTED || s->st_s[i] >= ST_INFECTED) s->infected.push_back(i);
}

/* weekday_generateAfterschoolAndErrandDestinations + sort + lower_bound
 * (PandemicSim.cu:3025-3044, :2972-3022, :1212-1248). */
void build_weekday_tables(or_sim* s) {
    s->errand_off = s->rand_offset;
    const int N = s->N;
    std::vector<int32_t> keys((size_t)2 * N), ids((size_t)2 * N);
    for (int g = 0; g <= N / 2; ++g) {
        const Words4 u = draw4(s->errand_off + (uint64_t)g, s->k0, s->k1);
        const int p1 = 2 * g;
        if (p1 < N) {
            weekday_dests(s, u.w[0], u.w[1], s->age[p1], &keys[p1], &keys[(size_t)N + p1]);
            ids[p1] = p1;
            ids[(size_t)N + p1] = p1;
        }
        const int p2 = p1 + 1;
        if (p2 < N) {
            weekday_dests(s, u.w[2], u.w[3], s->age[p2], &keys[p2], &keys[(size_t)N + p2]);
            ids[p2] = p2;
            ids[(size_t)N + p2] = p2;
        }
    }
    s->rand_offset += (uint64_t)(N / 2);
    bucket_stable(keys, ids, WD_HOURS * s->W, s->errand_offsets, s->errand_people);
    s->errand_hours_today = WD_HOURS;
}

/* weekend_assignErrands + sort + lower_bound (PandemicSim.cu:1404-1419,
 * :2082-2094, :1316-1354). */
void build_weekend_tables(or_sim* s) {
    s->errand_off = s->rand_offset;
    const int N = s->N;
    std::vector<int32_t> keys((size_t)3 * N), ids((size_t)3 * N);
    for (int p = 0; p < N; ++p) {
        const uint64_t off = s->errand_off + (uint64_t)(int64_t)(p * 2);
        weekend_errands(s, off, &keys[(size_t)3 * p]);
        ids[(size_t)3 * p] = ids[(size_t)3 * p + 1] = ids[(size_t)3 * p + 2] = p;
    }
    s->rand_offset += (uint64_t)(2 * (int64_t)N);
    bucket_stable(keys, ids, WE_HOURS * s->W, s->errand_offsets, s->errand_people);
    s->errand_hours_today = WE_HOURS;
}

/* Contact interventions (or_interventions) for one infector's chosen
 * contacts; the normalising contact count is not changed. */
void drop_contacts(const or_sim* s, bool weekend, uint64_t M, int C, int32_t* vic, uint8_t* typ) {
    const or_interventions& iv = s->iv;
    if (iv.school_close_days == 0 && iv.workplace_reduction == 0.f && iv.errand_reduction == 0.f) return;
    const Words4 x = draw4(M, s->ik0, s->ik1);
    const bool closed = !weekend && s->current_day >= iv.school_close_start &&
                        s->current_day < iv.school_close_start + iv.school_close_days;
    for (int c = 0; c < C; ++c) {
        const float u16 = (float)((x.w[c / 2] >> (16 * (c % 2))) & 0xFFFFu) * (1.0f / 65536.0f);
        bool drop = false;
        if (typ[c] == CT_SCHOOL || typ[c] == CT_AFTERSCHOOL)
            drop = closed;
        else if (typ[c] == CT_WORKPLACE)
            drop = u16 < iv.workplace_reduction;
        else if (typ[c] == CT_ERRAND)
            drop = u16 < iv.errand_reduction;
        if (drop) {
            vic[c] = -1;
            typ[c] = CT_NONE;
        }
    }
}

void contacts_day(or_sim* s, bool weekend) {
    const int I = (int)s->infected.size();
    const int C = weekend ? C_WE : C_WD;
    if (s->trace_on) {
        s->tr_infector.assign((size_t)I * C, -1);
        s->tr_victim.assign((size_t)I * C, -1);
        s->tr_type.assign((size_t)I * C, 0);
        s->tr_action.assign((size_t)I * C, 0);
        s->tr_draw.assign((size_t)I * C, 0);
        s->tr_prof.assign((size_t)I * C, 0);
    }
    int32_t vic[8];
    uint8_t typ[8];
    for (int pos = 0; pos < I; ++pos) {
        const int32_t idx = s->infected[pos];
        const uint64_t M = s->rand_offset + (uint64_t)(int64_t)(pos * 6);
        int kval;
        if (weekend)
            kval = make_contacts_weekend(s, idx, M, vic, typ);
        else
            kval = make_contacts_weekday(s, idx, s->age[idx], M, vic, typ);
        drop_contacts(s, weekend, M, C, vic, typ);
        process_contacts(s, idx, kval, vic, typ, C, M + 2, s->trace_on ? (int64_t)pos * C : -1);
    }
    s->rand_offset += (uint64_t)(6 * (int64_t)I);
}

}  // namespace

/* ============================ C API ============================ */
extern "C" {

const char* or_last_error(void) { return g_err.c_str(); }

void or_glibc_seeds(int32_t core_seed, int32_t out[4]) {
    srand((unsigned)core_seed);
    for (int i = 0; i < 4; ++i) out[i] = rand();
}

int or_glibc_initial_infected(int32_t core_seed, int32_t n_people, int32_t n, int32_t* out) {
    if (n < 0 || n > n_people) {
        g_err = "initial infected exceeds population";
        return -1;
    }
    srand((unsigned)core_seed);
    for (int i = 0; i < 4; ++i) (void)rand();
    for (int i = 0; i < n; ++i) {
        do {
            out[i] = rand() % n_people;
            for (int j = 0; j < i; ++j)
                if (out[j] == out[i]) {
                    out[i] = -1;
                    break;
                }
        } while (out[i] == -1);
    }
    return 0;
}

void or_threefry2x64_20(uint64_t c0, uint64_t c1, uint64_t k0, uint64_t k1, uint64_t out[2]) {
    tf2x64(c0, c1, k0, k1, out);
}
void or_threefry2x32_20(uint32_t c0, uint32_t c1, uint32_t k0, uint32_t k1, uint32_t out[2]) {
    tf2x32(c0, c1, k0, k1, out);
}

or_sim* or_create(const or_params* p, int32_t core_seed) {
    if (!p) return nullptr;
    if (p->max_days < 0 || p->max_days > 255) {
        g_err = "max_days must be in [0,255]";
        return nullptr;
    }
    or_sim* s = new or_sim();
    s->p = *p;
    s->core_seed = core_seed;
    or_glibc_seeds(core_seed, s->seed);
    s->k0 = (uint64_t)(int64_t)s->seed[0];
    s->k1 = (uint64_t)(int64_t)s->seed[1];
    s->ik0 = s->k0 ^ 0x9E3779B97F4A7C15ull; /* intervention stream key */
    s->ik1 = s->k1 ^ 0xBF58476D1CE4E5B9ull;
    scale_and_calibrate(s);
    return s;
}

void or_destroy(or_sim* s) { delete s; }

void or_set_interventions(or_sim* s, const or_interventions* iv) { s->iv = *iv; }

/* setupSim (PandemicSim.cu:146-208) */
int or_setup(or_sim* s) {
    const int H = s->H;
    /* K1/K2: household types, 4 per counter (PandemicSim.cu:3579-3605, :2329-2357) */
    s->hh_types.assign((size_t)H + 1, 0);
    int64_t npeople = 0;
    for (int g = 0; g <= H / 4; ++g) {
        const Words4 u = draw4((uint64_t)g, s->k0, s->k1);
        for (int i = 0; i < 4; ++i) {
            const int h = 4 * g + i;
            if (h < H) {
                const uint8_t t = fish_household(u.w[i]);
                s->hh_types[h] = t;
                npeople += HH_ADULTS[t] + HH_CHILDREN[t];
            }
        }
    }
    if (npeople > 0x7FFFFFFF / 8) {
        g_err = "population too large";
        return -1;
    }
    s->N = (int)npeople;
    const int N = s->N;
    if (s->p.init_pandemic < 0 || s->p.init_seasonal < 0 ||
        (int64_t)s->p.init_pandemic + s->p.init_seasonal > N) {
        g_err = "initial infected exceeds population";
        return -1;
    }
    s->rand_offset = 0;
    s->current_day = 0;
    s->rand_offset += (uint64_t)(H / 4);
    /* K3/K4 (PandemicSim.cu:2490-2531, :2579-2613) */
    s->hh_offsets.assign((size_t)H + 1, 0);
    s->household.assign((size_t)N, 0);
    s->age.assign((size_t)N, 0);
    int adults = 0, children = 0;
    std::vector<int> ascan((size_t)H + 1), cscan((size_t)H + 1);
    for (int h = 0; h <= H; ++h) {
        ascan[h] = adults;
        cscan[h] = children;
        if (h < H) {
            adults += HH_ADULTS[s->hh_types[h]];
            children += HH_CHILDREN[s->hh_types[h]];
        }
    }
    for (int h = 0; h < H; ++h) {
        const int t = s->hh_types[h];
        int off = ascan[h] + cscan[h];
        s->hh_offsets[h] = off;
        for (int k = 0; k < HH_ADULTS[t]; ++k) {
            s->household[off + k] = h;
            s->age[off + k] = AGE_ADULT;
        }
        off += HH_ADULTS[t];
        for (int k = 0; k < HH_CHILDREN[t]; ++k) {
            s->household[off + k] = h;
            s->age[off + k] = 0;
        }
    }
    s->hh_offsets[H] = N;
    /* K5 (PandemicSim.cu:3812-3830, :3750-3778) */
    s->workplace_off = s->rand_offset;
    s->wp_keys.assign((size_t)N, 0);
    for (int g = 0; g <= N / 4; ++g) {
        const Words4 u = draw4(s->workplace_off + (uint64_t)g, s->k0, s->k1);
        for (int i = 0; i < 4 && 4 * g + i < N; ++i) {
            const int idx = 4 * g + i;
            assign_workplace_or_school(s, u.w[i], &s->age[idx], &s->wp_keys[idx]);
        }
    }
    s->rand_offset += (uint64_t)(N / 4);
    /* K6: stable sort by workplace + lower_bound (PandemicSim.cu:3835-3862) */
    {
        std::vector<int32_t> ids((size_t)N);
        for (int i = 0; i < N; ++i) ids[i] = i;
        bucket_stable(s->wp_keys, ids, s->W, s->wp_offsets, s->wp_people);
    }
    /* K7 (PandemicSim.cu:3654-3664, :708-758) */
    s->st_p.assign((size_t)N, ST_SUSCEPTIBLE);
    s->st_s.assign((size_t)N, ST_SUSCEPTIBLE);
    s->day_p.assign((size_t)N, DAY_NONE);
    s->day_s.assign((size_t)N, DAY_NONE);
    s->gen_p.assign((size_t)N, GEN_NONE);
    s->gen_s.assign((size_t)N, GEN_NONE);
    if (s->iv.vaccination[0] > 0.f || s->iv.vaccination[1] > 0.f)
        for (int i = 0; i < N; ++i) {
            const Words4 x = draw4((1ull << 63) | (uint64_t)i, s->ik0, s->ik1);
            if (unit_float(x.w[0]) < s->iv.vaccination[0]) s->st_p[i] = ST_RECOVERED;
            if (unit_float(x.w[1]) < s->iv.vaccination[1]) s->st_s[i] = ST_RECOVERED;
        }
    const int ninit = s->p.init_pandemic + s->p.init_seasonal;
    std::vector<int32_t> init((size_t)ninit);
    if (or_glibc_initial_infected(s->core_seed, N, ninit, init.data()) != 0) return -1;
    for (int i = 0; i < ninit; ++i) {
        const int id = init[i];
        if (i < s->p.init_pandemic) {
            s->st_p[id] = ST_INFECTED;
            s->day_p[id] = 0;
            s->gen_p[id] = 0;
        } else {
            s->st_s[id] = ST_INFECTED;
            s->day_s[id] = 0;
            s->gen_s[id] = 0;
        }
    }
    s->setup_done = true;
    s->done = false;
    s->days_simulated = 0;
    return 0;
}

/* One iteration of runToCompletion (PandemicSim.cu:882-954). */
int or_step(or_sim* s, int32_t counts16[16], int32_t* infected_count) {
    if (!s->setup_done || s->done || s->current_day >= s->p.max_days) return 1;
    int32_t counts[16];
    count_and_recover(s, counts);
    if (counts16) std::memcpy(counts16, counts, sizeof(counts));
    s->days_simulated = s->current_day + 1;
    build_infected(s);
    const int I = (int)s->infected.size();
    if (infected_count) *infected_count = I;
    if (I == 0) {
        s->done = true;
        return 1;
    }
    const bool weekend = (s->current_day % 7) >= 5;
    if (weekend)
        build_weekend_tables(s);
    else
        build_weekday_tables(s);
    contacts_day(s, weekend);
    s->current_day++;
    if (s->current_day >= s->p.max_days) s->done = true;
    return 0;
}

int or_set_state(or_sim* s, int32_t day, uint64_t rand_offset, const uint8_t* st_p, const uint8_t* st_s,
                 const uint8_t* day_p, const uint8_t* day_s, const uint8_t* gen_p, const uint8_t* gen_s) {
    if (!s->setup_done || day < 0 || day > s->p.max_days) return -1;
    const size_t n = (size_t)s->N;
    std::memcpy(s->st_p.data(), st_p, n);
    std::memcpy(s->st_s.data(), st_s, n);
    std::memcpy(s->day_p.data(), day_p, n);
    std::memcpy(s->day_s.data(), day_s, n);
    std::memcpy(s->gen_p.data(), gen_p, n);
    std::memcpy(s->gen_s.data(), gen_s, n);
    s->current_day = day;
    s->days_simulated = day;
    s->rand_offset = rand_offset;
    s->done = day >= s->p.max_days;
    return 0;
}

int or_run(or_sim* s, int32_t* counts, int32_t* infected, int32_t cap_days) {
    int days = 0;
    while (days < cap_days) {
        int32_t c[16], I = 0;
        if (!s->setup_done || s->done || s->current_day >= s->p.max_days) break;
        or_step(s, c, &I);
        if (counts) std::memcpy(counts + 16 * days, c, sizeof(c));
        if (infected) infected[days] = I;
        days++;
    }
    return days;
}

int64_t or_get_size(or_sim* s, int which) {
    switch (which) {
        case OR_STATUS_P: case OR_STATUS_S: case OR_DAY_P: case OR_DAY_S:
        case OR_GEN_P: case OR_GEN_S: case OR_AGE: return s->N;
        case OR_HOUSEHOLD: return (int64_t)s->N * 4;
        case OR_HH_OFFSETS: return (int64_t)s->hh_offsets.size() * 4;
        case OR_WP_OFFSETS: return (int64_t)s->wp_offsets.size() * 4;
        case OR_WP_PEOPLE: return (int64_t)s->wp_people.size() * 4;
        case OR_INFECTED: return (int64_t)s->infected.size() * 4;
        case OR_ERRAND_OFFSETS: return (int64_t)s->errand_offsets.size() * 4;
        case OR_ERRAND_PEOPLE: return (int64_t)s->errand_people.size() * 4;
        case OR_HH_TYPES: return (int64_t)s->hh_types.size();
        case OR_WP_KEYS: return (int64_t)s->wp_keys.size() * 4;
    }
    return -1;
}

int or_get_array(or_sim* s, int which, void* dst, int64_t bytes) {
    const void* src = nullptr;
    switch (which) {
        case OR_STATUS_P: src = s->st_p.data(); break;
        case OR_STATUS_S: src = s->st_s.data(); break;
        case OR_DAY_P: src = s->day_p.data(); break;
        case OR_DAY_S: src = s->day_s.data(); break;
        case OR_GEN_P: src = s->gen_p.data(); break;
        case OR_GEN_S: src = s->gen_s.data(); break;
        case OR_AGE: src = s->age.data(); break;
        case OR_HOUSEHOLD: src = s->household.data(); break;
        case OR_HH_OFFSETS: src = s->hh_offsets.data(); break;
        case OR_WP_OFFSETS: src = s->wp_offsets.data(); break;
        case OR_WP_PEOPLE: src = s->wp_people.data(); break;
        case OR_INFECTED: src = s->infected.data(); break;
        case OR_ERRAND_OFFSETS: src = s->errand_offsets.data(); break;
        case OR_ERRAND_PEOPLE: src = s->errand_people.data(); break;
        case OR_HH_TYPES: src = s->hh_types.data(); break;
        case OR_WP_KEYS: src = s->wp_keys.data(); break;
        default: g_err = "bad array id"; return -1;
    }
    const int64_t size = or_get_size(s, which);
    if (bytes < size) {
        g_err = "destination too small";
        return -1;
    }
    if (size > 0) std::memcpy(dst, src, (size_t)size);
    return 0;
}

int64_t or_get_scalar(or_sim* s, int which) {
    switch (which) {
        case 0: return s->N;
        case 1: return s->H;
        case 2: return s->W;
        case 3: return (int64_t)s->rand_offset;
        case 4: return (int64_t)s->workplace_off;
        case 5: return (int64_t)s->errand_off;
        case 6: return s->current_day;
        case 7: return s->days_simulated;
        case 8: return s->offset[FIRST_ERRAND_ROW];
    }
    return -1;
}

int or_get_type_tables(or_sim* s, int32_t count[14], int32_t offset[14]) {
    for (int t = 0; t < NUM_TYPES; ++t) {
        count[t] = s->count[t];
        offset[t] = s->offset[t];
    }
    return 0;
}

int or_get_infection_tables(or_sim* s, float factor[2], float shed[60]) {
    factor[0] = s->factor[0];
    factor[1] = s->factor[1];
    std::memcpy(shed, s->shed, sizeof(s->shed));
    return 0;
}

/* final_countReproduction (PandemicSim.cu:2227-2304): histogram form of
 * sort + lower_bound. */
int or_generation_offsets(or_sim* s, int32_t* c_p, int32_t* c_s, int32_t len) {
    const int md = s->p.max_days;
    if (len < md + 1) return -1;
    int64_t hp[256] = {0}, hs[256] = {0};
    for (int i = 0; i < s->N; ++i) {
        hp[s->gen_p[i]]++;
        hs[s->gen_s[i]]++;
    }
    int64_t ap = 0, as = 0;
    for (int k = 0; k < md; ++k) {
        c_p[k] = (int32_t)ap;
        c_s[k] = (int32_t)as;
        ap += hp[k];
        as += hs[k];
    }
    c_p[md] = s->N;
    c_s[md] = s->N;
    return 0;
}

int32_t or_recalc_workplace(or_sim* s, int32_t idx, uint8_t age) { return recalc_workplace(s, idx, age); }
int32_t or_recalc_weekday(or_sim* s, int32_t idx, uint8_t age, int32_t dest[2]) {
    return recalc_weekday(s, idx, age, dest, dest + 1);
}
void or_recalc_weekend(or_sim* s, int32_t idx, int32_t dest[3]) { recalc_weekend(s, idx, dest); }
int or_get_raw_table(const char* name, double* out, int cap) {
    std::vector<double> v;
    const std::string n = name ? name : "";
    if (n == "wp_pdf") v.assign(WP_PDF, WP_PDF + NUM_TYPES);
    else if (n == "wd_pdf") v.assign(WD_PDF, WD_PDF + NUM_TYPES);
    else if (n == "we_pdf") v.assign(WE_PDF, WE_PDF + NUM_TYPES);
    else if (n == "wp_count") v.assign(WP_COUNT_BASE, WP_COUNT_BASE + NUM_TYPES);
    else if (n == "wp_max") v.assign(WP_MAX, WP_MAX + NUM_TYPES);
    else if (n == "hh_cdf") v.assign(HH_CDF, HH_CDF + HH_ROWS);
    else if (n == "hh_adults") v.assign(HH_ADULTS, HH_ADULTS + HH_ROWS);
    else if (n == "hh_children") v.assign(HH_CHILDREN, HH_CHILDREN + HH_ROWS);
    else if (n == "child_cdf") v.assign(CHILD_CDF, CHILD_CDF + CHILD_ROWS);
    else if (n == "wd_assign") v.assign(&WD_ASSIGN[0][0], &WD_ASSIGN[0][0] + 8);
    else if (n == "we_assign") v.assign(&WE_ASSIGN[0][0], &WE_ASSIGN[0][0] + 12);
    else if (n == "shed") v.assign(&SHED_BASE[0][0], &SHED_BASE[0][0] + 60);
    else return -1;
    if ((int)v.size() > cap) return -1;
    for (size_t i = 0; i < v.size(); ++i) out[i] = v[i];
    return (int)v.size();
}

void or_set_offsets(or_sim* s, uint64_t workplace_off, uint64_t errand_off) {
    s->workplace_off = workplace_off;
    s->errand_off = errand_off;
}
int32_t or_fish_workplace(or_sim* s, uint32_t r) { return fish_workplace(s, r); }
void or_fish_school_and_age(or_sim* s, uint32_t r, uint8_t* age, int32_t* school) {
    fish_school_and_age(s, r, age, school);
}
uint8_t or_fish_household_type(uint32_t r) { return fish_household(r); }
int32_t or_fish_weekday_dest(or_sim* s, uint32_t r, uint8_t age) { return fish_weekday_dest(s, r, age); }
int32_t or_fish_weekend_dest(or_sim* s, uint32_t r) { return fish_weekend_dest(s, r); }

void or_enable_trace(or_sim* s, int on) { s->trace_on = on != 0; }
int64_t or_trace_size(or_sim* s) { return (int64_t)s->tr_victim.size(); }
int or_get_trace(or_sim* s, int32_t* infector, int32_t* victim, uint8_t* type, uint8_t* action) {
    const size_t n = s->tr_victim.size();
    if (n == 0) return 0;
    std::memcpy(infector, s->tr_infector.data(), n * 4);
    std::memcpy(victim, s->tr_victim.data(), n * 4);
    std::memcpy(type, s->tr_type.data(), n);
    std::memcpy(action, s->tr_action.data(), n);
    return 0;
}
int or_get_trace_draws(or_sim* s, uint8_t* draw, uint8_t* profile) {
    const size_t n = s->tr_draw.size();
    if (n == 0) return 0;
    if (draw) std::memcpy(draw, s->tr_draw.data(), n);
    if (profile) std::memcpy(profile, s->tr_prof.data(), n);
    return 0;
}
uint8_t or_infection_profile(or_sim* s, uint32_t r) { return infection_profile(s, r); }
float or_infection_probability(or_sim* s, int profile, int dd, int strain, int kval_sum) {
    return infection_probability(s, profile, dd, strain, kval_sum);
}
int or_plant(void) { return OR_PLANT; }

/* Output files: output_infected_stats.csv (PandemicSim.cu:263, :2771-2818),
 * infected_max.txt (:855-866, :923-931), output_rn.csv (:2276-2297). */
int or_write_outputs(or_sim* s, const char* dir, const int32_t* counts, const int32_t* infected, int32_t days) {
    const std::string d = dir ? std::string(dir) + "/" : std::string();
    FILE* f = fopen((d + "output_infected_stats.csv").c_str(), "w");
    if (!f) return -1;
    fprintf(f, "day,pandemic_susceptible,pandemic_infectious,pandemic_symptomatic,pandemic_asymptomatic,"
               "pandemic_recovered,seasonal_susceptible,seasonal_infectious,seasonal_symptomatic,"
               "seasonal_asymptomatic,seasonal_recovered\n");
    float max_prop = 0.f;
    int max_day = 0;
    for (int day = 0; day < days; ++day) {
        const int32_t* c = counts + 16 * day;
        const int psym = c[2] + c[3] + c[4], pasym = c[5] + c[6] + c[7];
        const int ssym = c[10] + c[11] + c[12], sasym = c[13] + c[14] + c[15];
        fprintf(f, "%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d\n", day, c[1], psym + pasym, psym, pasym, c[0], c[9],
                ssym + sasym, ssym, sasym, c[8]);
        if (infected[day] > 0) {
            volatile float prop = (float)infected[day] / s->N;
            if (prop > max_prop) {
                max_prop = prop;
                max_day = day;
            }
        }
    }
    fclose(f);
    f = fopen((d + "infected_max.txt").c_str(), "w");
    if (!f) return -1;
    fprintf(f, "max_proportion,day\n%f,%d\n", max_prop, max_day);
    fclose(f);
    const int md = s->p.max_days;
    std::vector<int32_t> cp((size_t)md + 1), cs((size_t)md + 1);
    or_generation_offsets(s, cp.data(), cs.data(), md + 1);
    f = fopen((d + "output_rn.csv").c_str(), "w");
    if (!f) return -1;
    fprintf(f, "gen,gen_size_p,rn_p,gen_size_s,rn_s\n");
    for (int g = 0; g < md - 1; ++g) {
        const int sp = cp[g + 1] - cp[g], np = cp[g + 2] - cp[g + 1];
        volatile float rp = (float)np / sp;
        const int ss = cs[g + 1] - cs[g], ns = cs[g + 2] - cs[g + 1];
        volatile float rs = (float)ns / ss;
        fprintf(f, "%d,%d,%f,%d,%f\n", g, sp, (double)rp, ss, (double)rs);
    }
    fclose(f);
    return 0;
}

}  // extern "C"
