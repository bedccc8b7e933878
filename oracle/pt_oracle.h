/*
 * CPU ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * Sequential CPU restatement of the PandemicThrust daily simulation loop
 * (reference: PandemicThrust/PandemicThrust/PandemicSim.cu).  It exists to
 * check the HIP implementation; it is never linked into, loaded by or called
 * from the product library.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use it.
 *
 * Semantics: the reference applies infections with a racy atomicCAS
 * (PandemicSim.cu:3048-3067).  The oracle applies them in the canonical
 * sequential order: infectors by ascending infected-list position, contacts
 * by slot, pandemic before seasonal; an attempt succeeds iff its draw
 * succeeds and the victim is still susceptible for that strain.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* constants.csv fields in file order (PandemicSim.cu:299-309). */
typedef struct or_params {
    int32_t max_days;
    float r_pandemic;
    float r_seasonal;
    int32_t init_pandemic;
    int32_t init_seasonal;
    float people_scaling;
    float location_scaling;
    float percent_symptomatic;
    float asymp_factor;
} or_params;

/* Interventions (not in the reference; a build extension).  Neutral:
 * latent {0,0}, infectious {10,10}, no closure, no reductions, no
 * vaccination — the reference exactly.  Infectious while latent <= day -
 * infection day < latent + infectious (shedding index day - infection day -
 * latent), recovery at latent + infectious days.  On weekdays inside
 * [school_close_start, +school_close_days) school and after-school contacts
 * are dropped; a workplace (errand) contact in slot c is dropped when
 * uniform16(c) < workplace (errand) reduction, uniform16(c) being 16-bit
 * field c of Threefry2x64(ctr = the infector's contact counter, key = SEED
 * XOR (0x9E3779B97F4A7C15, 0xBF58476D1CE4E5B9)); dropped slots become
 * (-1, NONE) and the infection probability keeps the count of contacts chosen
 * before the drops.  Agent i is immune to strain k from setup when word k of
 * the same stream at counter 2^63 | i, over 2^32, is below vaccination[k]. */
typedef struct or_interventions {
    int32_t latent_days[2];
    int32_t infectious_days[2];
    int32_t school_close_start;
    int32_t school_close_days;
    float workplace_reduction;
    float errand_reduction;
    float vaccination[2];
} or_interventions;

typedef struct or_sim or_sim;

/* Array ids for or_get_array. */
enum {
    OR_STATUS_P = 0, OR_STATUS_S, OR_DAY_P, OR_DAY_S, OR_GEN_P, OR_GEN_S, /* uint8[N] */
    OR_AGE,                 /* uint8[N] */
    OR_HOUSEHOLD,           /* int32[N] */
    OR_HH_OFFSETS,          /* int32[H+1] */
    OR_WP_OFFSETS,          /* int32[W+1] */
    OR_WP_PEOPLE,           /* int32[N] */
    OR_INFECTED,            /* int32[I] (today's list) */
    OR_ERRAND_OFFSETS,      /* int32[hours*W+1] (today's table) */
    OR_ERRAND_PEOPLE,       /* int32[E*N] (today's table) */
    OR_HH_TYPES,            /* uint8[H+1] */
    OR_WP_KEYS              /* int32[N] workplace/school of each person from setup */
};

/* glibc seed sequence: srand(core_seed); SEED[i] = rand(), i < 4
 * (PandemicSim.cu:2853-2861). */
void or_glibc_seeds(int32_t core_seed, int32_t out[4]);
/* srand(core_seed), skip 4 draws, then n unique rand() % n_people
 * (PandemicSim.cu:1642-1659). */
int or_glibc_initial_infected(int32_t core_seed, int32_t n_people, int32_t n, int32_t* out);

/* Random123 Threefry, written from the published algorithm. */
void or_threefry2x64_20(uint64_t c0, uint64_t c1, uint64_t k0, uint64_t k1, uint64_t out[2]);
void or_threefry2x32_20(uint32_t c0, uint32_t c1, uint32_t k0, uint32_t k1, uint32_t out[2]);

or_sim* or_create(const or_params* p, int32_t core_seed);
void or_destroy(or_sim* s);
const char* or_last_error(void);

/* Setup (K1-K7).  Returns 0 on success. */
int or_setup(or_sim* s);
/* Interventions for the next or_setup (default: none). */
void or_set_interventions(or_sim* s, const or_interventions* iv);

/* One reference loop iteration.  Writes the day's 16 status counts and I.
 * Returns 0 if the day ran contacts, 1 if it was the terminal I==0 day or
 * max_days is reached (then counts are not written past max_days). */
int or_step(or_sim* s, int32_t counts16[16], int32_t* infected_count);
/* Continues a simulation from a saved day boundary (after or_setup): the six
 * uint8 per-agent arrays, the day about to run and the running RNG offset
 * (what a run reaches after `day` days).  Returns 0, or -1 before setup. */
int or_set_state(or_sim* s, int32_t day, uint64_t rand_offset, const uint8_t* st_p, const uint8_t* st_s,
                 const uint8_t* day_p, const uint8_t* day_s, const uint8_t* gen_p, const uint8_t* gen_s);
/* Runs the remaining days; counts[days*16], infected[days]; returns days run. */
int or_run(or_sim* s, int32_t* counts, int32_t* infected, int32_t cap_days);

int64_t or_get_size(or_sim* s, int which);
int or_get_array(or_sim* s, int which, void* dst, int64_t bytes);

/* Scalars: 0 N, 1 H, 2 W, 3 rand_offset, 4 workplace_off, 5 errand_off,
 * 6 current_day, 7 days_simulated (rows written), 8 errand locations base. */
int64_t or_get_scalar(or_sim* s, int which);

/* Scaled per-type tables after setup_scaleSimulation (PandemicSim.cu:1127). */
int or_get_type_tables(or_sim* s, int32_t count[14], int32_t offset[14]);
/* Infectiousness factors and the shedding table after
 * setup_calculateInfectionData (PandemicSim.cu:2828). */
int or_get_infection_tables(or_sim* s, float factor[2], float shed[60]);

/* Generation histograms: c[k] = #agents with gen < k for k < max_days,
 * c[max_days] = N (PandemicSim.cu:2236-2260).  len = max_days+1. */
int or_generation_offsets(or_sim* s, int32_t* c_p, int32_t* c_s, int32_t len);

/* Per-person regeneration hooks (the reference's host twins,
 * host_functions.cpp).  Valid after setup / after the day's errands. */
int32_t or_recalc_workplace(or_sim* s, int32_t idx, uint8_t age);
int32_t or_recalc_weekday(or_sim* s, int32_t idx, uint8_t age, int32_t dest[2]);
void or_recalc_weekend(or_sim* s, int32_t idx, int32_t dest[3]);

/* Raw model tables by name (wp_pdf, wd_pdf, we_pdf, wp_count, wp_max,
 * hh_cdf, hh_adults, hh_children, child_cdf, wd_assign, we_assign, shed),
 * row-major, as doubles.  Returns the element count or -1. */
int or_get_raw_table(const char* name, double* out, int cap);

/* Test hook: override the saved RNG offsets used by the regeneration hooks. */
void or_set_offsets(or_sim* s, uint64_t workplace_off, uint64_t errand_off);

/* Stand-alone draw functions with explicit tables (golden-vector hooks). */
int32_t or_fish_workplace(or_sim* s, uint32_t r);
void or_fish_school_and_age(or_sim* s, uint32_t r, uint8_t* age, int32_t* school);
uint8_t or_fish_household_type(uint32_t r);
int32_t or_fish_weekday_dest(or_sim* s, uint32_t r, uint8_t age);
int32_t or_fish_weekend_dest(or_sim* s, uint32_t r);

/* Per-contact trace of the last contact day (the SIM_VALIDATION layout,
 * PandemicSim.cu:3221-3230): per slot infector, victim, type, action. */
int64_t or_trace_size(or_sim* s);
int or_get_trace(or_sim* s, int32_t* infector, int32_t* victim, uint8_t* type, uint8_t* action);
void or_enable_trace(or_sim* s, int on);

/* Draws of the same slots: draw bit 1 / bit 0 = the pandemic / seasonal
 * draw y < p * kval succeeded (PandemicSim.cu:3195-3206) whatever the
 * victim's state; profile = the profile each draw gives (pandemic low
 * nibble, seasonal high nibble).  Either pointer may be NULL. */
int or_get_trace_draws(or_sim* s, uint8_t* draw, uint8_t* profile);

/* device_getInfectionProfile (PandemicSim.cu:3974-3988) and
 * device_calculateInfectionProbability (:2307-2326) with the simulation's
 * calibrated tables (valid after or_create). */
uint8_t or_infection_profile(or_sim* s, uint32_t r);
float or_infection_probability(or_sim* s, int profile, int dd, int strain, int kval_sum);
/* Which deliberate mistake this compiled oracle carries (0: none; see
 * pt_oracle.cpp). */
int or_plant(void);

/* Writes the reference's output files into dir. */
int or_write_outputs(or_sim* s, const char* dir, const int32_t* counts, const int32_t* infected, int32_t days);

#ifdef __cplusplus
}
#endif
