/*
 * pt_abi.h — C ABI of the MI355X (gfx950) PandemicThrust daily-loop engine.
 *
 * The reference has no plugin or FFI interface: its daily loop sits behind
 * the PandemicSim class (PandemicSim.h:68-281) whose device state lives in
 * process-global __constant__ symbols (PandemicSim.cu:26-96).  This header
 * replaces that kernel-level boundary and the device half of the class with
 * plain C entry points: one handle per simulation, no global device state,
 * integer return codes (0 ok, negative error; the library never exits), and
 * caller-owned host buffers.  A handle binds one HIP stream and is not
 * thread-safe; several handles may coexist in one process.
 *
 * Citations are relative to PandemicThrust/PandemicThrust/ of the reference.
 */
#ifndef PANDEMIC_PT_ABI_H
#define PANDEMIC_PT_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PT_OK 0
#define PT_EINVAL (-1)   /* bad argument or input outside the model's range */
#define PT_ENOMEM (-2)   /* device allocation failed */
#define PT_EDEVICE (-3)  /* HIP runtime error */
#define PT_ESTATE (-4)   /* call out of order (e.g. step before setup) */

/* The data row of constants.csv in file order (PandemicSim.cu:299-309). */
typedef struct pt_params {
    int32_t max_days;            /* 1..255 (day_t is 8 bits, PandemicSim.h:76) */
    float r_pandemic;
    float r_seasonal;
    int32_t init_pandemic;
    int32_t init_seasonal;
    float people_scaling;        /* households = round(people_scaling * 1e6) */
    float location_scaling;      /* per-type workplace counts scaled by this */
    float percent_symptomatic;
    float asymp_factor;
} pt_params;

/* Interventions.  The reference has none; the thesis sketches closures and
 * vaccination as future work (thesis.pdf p.18-19).  The neutral set
 * {0,0}, {10,10}, 0, 0, 0, 0, {0,0} reproduces the reference exactly.
 * Contact drops and vaccination draw from their own Threefry stream, so they
 * consume no draw of the model's stream. */
typedef struct pt_interventions {
    int32_t latent_days[2];      /* per strain, 0..30: infectious from this many days after infection */
    int32_t infectious_days[2];  /* per strain, 1..10: length of the infectious window (shedding
                                    profile truncated), recovery after latent + infectious days */
    int32_t school_close_start;  /* first day of a weekday school closure */
    int32_t school_close_days;   /* closure length (0: schools stay open); closed schools drop
                                    school and after-school contacts */
    float workplace_reduction;   /* 0..1: share of adults' workplace contacts dropped */
    float errand_reduction;      /* 0..1: share of errand contacts dropped */
    float vaccination[2];        /* 0..1: share of agents immune to each strain from setup */
} pt_interventions;

typedef struct pt_sim pt_sim;

/* Text of the last error raised on the calling thread. */
const char* pt_last_error(void);

/* glibc seed sequence of setup_loadSeed (PandemicSim.cu:2843-2862):
 * srand(core_seed); seeds[i] = rand() for i < 4.  Host only. */
int pt_glibc_seeds(int32_t core_seed, int32_t seeds_out[4]);

/* Initial infected of setup_initialInfected (PandemicSim.cu:708-719) via
 * n_unique_numbers (:1642-1659): srand(core_seed), skip the 4 seed draws,
 * then n unique rand() % n_people.  Host only. */
int pt_glibc_initial_infected(int32_t core_seed, int32_t n_people, int32_t n, int32_t* ids_out);

/* PandemicSim() constructor device half (PandemicSim.cu:98-133):
 * setup_scaleSimulation (:1127), setup_calculateInfectionData (:2828) and
 * setup_pushDeviceData (:569).  seed_key = SEED[0], SEED[1] (the Threefry
 * key).  device = HIP device ordinal. */
int pt_sim_create(const pt_params* params, const int32_t seed_key[2], int32_t device, pt_sim** out);

/* Sets the interventions of a simulation (no counterpart in the reference;
 * for scenario sweeps).  Call before pt_sim_setup; default: none. */
int pt_sim_set_interventions(pt_sim* sim, const pt_interventions* iv);

/* setupSim (PandemicSim.cu:146-208) up to, not including, the initial
 * infections: population size, households, workplaces, workplace table and
 * status initialisation (K1-K7).  Returns the population size. */
int pt_sim_setup(pt_sim* sim, int64_t* n_people_out);

/* setup_initialInfected (PandemicSim.cu:708-758): the first n_pandemic ids
 * get the pandemic strain, the next n_seasonal the seasonal strain
 * (status 2, day 0, generation 0).  Ids must be unique and < n_people. */
int pt_sim_seed_infections(pt_sim* sim, const int32_t* ids, int32_t n_pandemic, int32_t n_seasonal);

/* One iteration of runToCompletion (PandemicSim.cu:882-954): count and
 * recover, compact, then the weekday or weekend step.  Writes the day's 16
 * status counts (daily_writeInfectedStats order: pandemic codes 0..7, then
 * seasonal 0..7) and the infected count.  Returns 0 after a day with
 * contacts, 1 for the terminal I == 0 day (its counts are still reported)
 * or when the loop had already ended (nothing is written), negative on
 * error. */
int pt_sim_step(pt_sim* sim, int32_t counts16_out[16], int32_t* infected_out);

/* Runs the remaining days of runToCompletion without per-day host
 * synchronisation.  counts_out[days*16], infected_out[days] (either may be
 * NULL); *days_out = number of rows (days whose counts were produced). */
int pt_sim_run(pt_sim* sim, int32_t* counts_out, int32_t* infected_out, int32_t* days_out);

/* final_countReproduction (PandemicSim.cu:2227-2304) in histogram form:
 * c[k] = number of agents whose generation is < k for k < max_days, and
 * c[max_days] = n_people.  len must be >= max_days + 1. */
int pt_sim_generation_offsets(pt_sim* sim, int32_t* c_pandemic, int32_t* c_seasonal, int32_t len);

/* Parity hooks: copy a device array into a host buffer. */
enum pt_array {
    PT_STATUS_P = 0, PT_STATUS_S, PT_DAY_P, PT_DAY_S, PT_GEN_P, PT_GEN_S, /* uint8[N] */
    PT_AGE,             /* uint8[N] */
    PT_HOUSEHOLD,       /* int32[N] people_households */
    PT_HH_OFFSETS,      /* int32[H+1] household_offsets */
    PT_WP_OFFSETS,      /* int32[W+1] workplace_offsets */
    PT_WP_PEOPLE,       /* int32[N] workplace_people */
    PT_INFECTED,        /* int32[I] infected_indexes of the last day run */
    PT_ERRAND_OFFSETS,  /* int32[hours*W+1] errand_locationOffsets of the last day, reference layout */
    PT_ERRAND_PEOPLE,   /* int32[E*N] sorted errand table of the last day */
    PT_HH_TYPES,        /* uint8[H+1] household types */
    PT_WORKPLACE,       /* int32[N] workplace or school of each agent (validation; people_workplaces) */
    PT_ERRANDS,         /* int32[E*N] the last day's errand destinations as hour * W + location
                           (validation; people_errands: weekday hour-major, weekend 3 per person) */
    PT_LOCATION_TYPE_OFFSETS /* int32[15] first location of each of the 14 location types, then W
                                (WORKPLACE_TYPE_OFFSET_HOST, the scan of PandemicSim.cu:1155-1158) */
};
int64_t pt_sim_array_bytes(pt_sim* sim, int32_t which);
int pt_sim_copy_array(pt_sim* sim, int32_t which, void* dst, int64_t bytes);

/* Scalars: [0] N, [1] H, [2] W, [3] rand_offset, [4] workplace rand offset,
 * [5] errand rand offset of the last day, [6] current day, [7] days run. */
int pt_sim_get_scalars(pt_sim* sim, int64_t out[8]);

/* Errand-table layout: [0] weekday base location, [1] weekday hour-0
 * locations E_wd, [2] weekend base location, [3] weekend locations per hour
 * E_we, [4] weekday buckets, [5] weekend buckets (10 E_we), [6] radix passes
 * of a weekday table, [7] radix passes of a weekend table.  Weekend bucket
 * (hour, location) is stored at hour * E_we + (location - base).  Weekday
 * hour 0 is stored at location - base; weekday hour 1 stores only the
 * locations adults can reach (children stay at their after-school site and
 * make no hour-1 contacts), the rest share one last, never-queried bucket.
 * The PT_ERRAND_* hooks always return the reference's layout. */
int pt_sim_get_layout(pt_sim* sim, int64_t out[8]);

/* Errand-table strategy: 0 automatic (default), 1 build every bucket each
 * day, 2 build only the buckets that day's infectors draw contacts from.
 * Simulation results are identical in all modes; the PT_ERRAND_* parity
 * hooks show complete tables only in mode 1. */
int pt_sim_set_table_mode(pt_sim* sim, int32_t mode);

/* Measurement hook: the errand table of each day run since setup / restore.
 * full_out[d] = 1 if day d built every bucket, 0 if only the queried ones;
 * pairs_out[d] = pairs sorted that day.  Fills min(len, days run) entries
 * and returns their number (either pointer may be NULL). */
int pt_sim_table_log(pt_sim* sim, int32_t* full_out, int32_t* pairs_out, int32_t len);

/* Saves the post-setup, post-seeding state on the device; restore brings
 * the simulation back to it (used to repeat the daily loop). */
int pt_sim_snapshot(pt_sim* sim);
int pt_sim_restore(pt_sim* sim);

/* Per-kernel-class timing with HIP events carried in the kernels' dispatch
 * packets on the handle's stream (no extra barrier packets).  enable: 0 off,
 * negative = every class, k + 1 = class k only; any call clears the totals.
 * pt_sim_kernel_times fills ms_out[k] (total milliseconds) and
 * launches_out[k] for k < n and returns the number of classes;
 * pt_sim_kernel_name names class k. */
int pt_sim_enable_timing(pt_sim* sim, int32_t enable);
int pt_sim_kernel_times(pt_sim* sim, double* ms_out, int64_t* launches_out, int32_t n);
const char* pt_sim_kernel_name(int32_t k);

/* Per-day kernel-class times (profiler.cu's per-function rows of each day,
 * PandemicSim.cu:1186-1420, :2207-2224, :2738-2825): ms_out[d * C + k] =
 * milliseconds of class k on day d, C = the class count pt_sim_kernel_times
 * returns.  Covers the days enqueued since timing was last enabled; fills
 * min(max_days, days) rows and returns days (ms_out may be NULL). */
int pt_sim_day_kernel_times(pt_sim* sim, double* ms_out, int32_t max_days);

/* Setup phases, timed by every handle (milliseconds), after the reference's
 * profiled setup functions. */
enum pt_setup_phase {
    PT_SETUP_SCALE = 0,         /* setup_scaleSimulation :1127-1162 + setup_calculateInfectionData :2828 (host) */
    PT_SETUP_PUSH,              /* setup_pushDeviceData :569-700: handle, stream, model buffers */
    PT_SETUP_POPULATION_SIZE,   /* setup_calcPopulationSize_thrust :3634-3651 (K1/K2) */
    PT_SETUP_SIZE_ARRAYS,       /* setup_sizeGlobalArrays :1047-1107 (allocation) */
    PT_SETUP_HOUSEHOLDS,        /* setup_generateHouseholds :2552-2620 (K3/K4) */
    PT_SETUP_WORKPLACES,        /* setup_assignWorkplaces :3812-3868 (K5/K6) */
    PT_SETUP_STATUS,            /* setup_initializeStatusArrays :3654-3664 (K7) */
    PT_SETUP_INITIAL_INFECTED,  /* setup_initialInfected :708-758 (seeding) */
    PT_SETUP_PHASES
};
int pt_sim_setup_times(pt_sim* sim, double* ms_out, int32_t n);

/* Every daily row since setup (the rows of output_infected_stats.csv,
 * PandemicSim.cu:2771-2818, including rows restored from a checkpoint):
 * counts_out[rows * 16] and infected_out[rows] (either may be NULL).
 * Returns the number of rows copied, min(max_rows, rows). */
int pt_sim_daily_rows(pt_sim* sim, int32_t* counts_out, int32_t* infected_out, int32_t max_rows);

/* Checkpoint / resume at a day boundary (no counterpart in the reference,
 * whose run is one process from setupSim to the end of runToCompletion).
 * save writes pt_sim_checkpoint_bytes bytes into a caller buffer: the
 * per-agent status/day/generation arrays, the loop counters (current day,
 * rand_offset) and the daily rows so far, with a checksum.  load requires a
 * handle created and set up from the same constants, seed and interventions
 * (checked; the population is rebuilt by pt_sim_setup, not stored) and
 * replaces seeding: the loop then continues exactly where the saved one
 * stood, with results identical to an uninterrupted run. */
int64_t pt_sim_checkpoint_bytes(pt_sim* sim);
int pt_sim_checkpoint_save(pt_sim* sim, void* buf, int64_t bytes);
int pt_sim_checkpoint_load(pt_sim* sim, const void* buf, int64_t bytes);

/* Device memory and name for output_resource_log.csv / output_mem.csv
 * (cudaMemGetInfo, resource_logging.cu:37, :64). */
int pt_device_memory(int32_t device, uint64_t* free_bytes, uint64_t* total_bytes);
int pt_device_name(int32_t device, char* buf, int32_t len);

/* Validation trace (SIM_VALIDATION, PandemicSim.h:50-58, debug.cu:138-442,
 * :847-950): with tracing on, each contact day records per infector
 * position and slot (C = 8 weekday / 5 weekend) the victim, the contact type
 * (CONTACT_TYPE_*, simParameters.h:59-64), the contact's location
 * (household, workplace or school, after-school location, errands as
 * hour * W + location, -1 for none) and the action (ACTION_INFECT_*,
 * simParameters.h:23-26: 2 pandemic, 1 seasonal, 3 both, for the attempts
 * that infected the victim in the canonical order).  Costs 12 bytes per slot
 * of device memory (8N slots); results are unchanged.  pt_sim_trace copies
 * the last day's pt_sim_trace_size entries; infector = the infected agent of
 * the slot's position.  Returns the number of entries. */
int pt_sim_enable_trace(pt_sim* sim, int32_t on);
int64_t pt_sim_trace_size(pt_sim* sim);
int pt_sim_trace(pt_sim* sim, int32_t* infector, int32_t* victim, uint8_t* type, int32_t* loc, uint8_t* action,
                 int64_t n);
/* The draws of the same slots, as the reference's validation mode records
 * its uniforms and thresholds (rand_arr_1..4, PandemicSim.cu:3220-3230):
 * draw bit 1 / bit 0 = the pandemic / seasonal draw succeeded (y < p * kval,
 * :3195-3206), whatever the victim's state; profile = the infection profile
 * each draw gives (device_getInfectionProfile, :3974-3988; pandemic in the
 * low nibble, seasonal in the high nibble), computed for every slot as the
 * reference does.  Either pointer may be NULL.  Returns the number of
 * entries (pt_sim_trace_size). */
int pt_sim_trace_draws(pt_sim* sim, uint8_t* draw, uint8_t* profile, int64_t n);

/* Launch shapes and table-strategy thresholds.  They change only speed:
 * results are identical for every accepted value (the kernels stride over
 * their work, the radix sort is stable at any tile size, and a table of the
 * queried buckets answers every query a full table does).
 * pandemicthrust_amd/autotune.py picks them by measurement. */
enum pt_launch_knob {
    PT_LAUNCH_CONTACT_BLOCKS_PER_CU = 0, /* blocks per compute unit of the mark, contact and apply kernels
                                            (1..64, default 8) */
    PT_LAUNCH_GEN_BLOCKS_PER_CU,         /* blocks per compute unit of the errand-key kernels (1..64, default 16) */
    PT_LAUNCH_SORT_WIDE_BITS,            /* radix passes of at least this many bits (at most 8) run 1024-thread
                                            blocks on 8192-pair tiles, narrower ones 512 on 4096 (1..9,
                                            default 8; 9 = never) */
    PT_LAUNCH_MARK_LIMIT,                /* the day's queried buckets are marked (for a possible subset table)
                                            while infected <= K * value / 100, K = buckets (0..1000, default 100) */
    PT_LAUNCH_SUBSET_SHARE,              /* a marked day builds a subset table while marked buckets <
                                            K * value / 100 (0..100, default 75) */
    PT_LAUNCH_SORT_MAX_BITS,             /* widest digit of a radix pass: K buckets take ceil(log2 K / value)
                                            passes (4..11; default 8, 9 in a replicate batch) */
    PT_LAUNCH_SEGMENTED_TABLES,          /* replicate batch: 1 = each replicate's day table sorted in its own
                                            segment over its local keys (batches of up to 1024 replicates),
                                            0 = one sort over every replicate's key range (0..1, default 1;
                                            a single simulation ignores it) */
    PT_LAUNCH_KNOBS
};
int pt_sim_set_launch(pt_sim* sim, int32_t knob, int32_t value);
int pt_sim_get_launch(pt_sim* sim, int32_t knob);

/* Location table on caller-owned device memory: the reference's daily
 * cub::DeviceRadixSort::SortPairs + thrust::lower_bound
 * (PandemicSim.cu:1212-1248, :1316-1354; workplace table :3835-3862) as one
 * call.  keys[n]: uint32 bucket keys, each < n_buckets, of pairs whose ids
 * are their positions 0..n-1.  Writes people[n] = the ids grouped by key,
 * ascending within a key (the stable sort's order), and
 * offsets[n_buckets + 1] with offsets[k] = number of keys below k (so
 * offsets[n_buckets] = n).  All three are device pointers on `device`;
 * stream is a hipStream_t (NULL: the default stream).  The radix scratch
 * (about 20 bytes per pair) is allocated for the call and freed before it
 * returns, so the call synchronises the stream.  n < 2^31, n_buckets >= 1.
 * Keys are not checked on the device: a key >= n_buckets leaves the table
 * unspecified, but every write stays inside people[n] and
 * offsets[n_buckets + 1]. */
int pt_location_table(const uint32_t* keys, int64_t n, int32_t n_buckets, int32_t* people, int32_t* offsets,
                      int32_t device, void* stream);

/* The same table, allocating nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The radix
 * scratch is a file-scope device array of the library, one copy per device
 * (about 85 MB), holding at most PT_LOCATION_TABLE_NOALLOC_MAX_PAIRS pairs;
 * larger n is refused with PT_EINVAL.  The call only enqueues work on
 * stream and returns; results are identical to pt_location_table's.  Every
 * call on one device shares that device's array, so calls that overlap on
 * one device (other streams or threads) race: serialise them. */
#define PT_LOCATION_TABLE_NOALLOC_MAX_PAIRS 4194304
int pt_location_table_noalloc(const uint32_t* keys, int64_t n, int32_t n_buckets, int32_t* people,
                              int32_t* offsets, int32_t device, void* stream);

/* Blocks until the handle's stream is idle; loop time of the last
 * pt_sim_run in milliseconds (HIP events around the enqueued days). */
int pt_sim_synchronize(pt_sim* sim);
double pt_sim_last_run_ms(pt_sim* sim);

void pt_sim_destroy(pt_sim* sim);

#ifdef __cplusplus
}
#endif

#endif /* PANDEMIC_PT_ABI_H */
