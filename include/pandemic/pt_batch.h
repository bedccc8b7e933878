/*
 * pt_batch.h — C ABI of the replicate-batch engine: many independent
 * simulations (replicates) advanced in lockstep on one GPU.
 *
 * The reference runs one simulation per process and batches by running the
 * executable once per constants.csv row (PythonBatcher/write_gpu_constants.py:6-49
 * writes the rows; main.cu:12-31 is one run).  A batch replaces that loop:
 * replicate r has its own constants row and Threefry key, and ends exactly
 * where a pt_sim run (pt_abi.h) with the same inputs ends — the same daily
 * counts, the same per-agent state.  All replicates share every kernel
 * launch: their agents, households, workplaces and errand buckets lie side
 * by side in HBM, and each day's errand table of every replicate is built by
 * one sort.
 *
 * Conventions as in pt_abi.h: 0 ok, negative error (PT_E*), pt_last_error
 * holds the text; one handle binds one HIP stream and is not thread-safe.
 */
#ifndef PANDEMIC_PT_BATCH_H
#define PANDEMIC_PT_BATCH_H

#include <stdint.h>

#include "pt_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pt_batch pt_batch;

/* Capacity: replicate r occupies its population rounded up to a multiple
 * of PT_BATCH_SLOT_TILE agent slots, and a batch holds at most
 * PT_BATCH_MAX_SLOTS slots in all (agent ids are int32 and a weekend's 3 pairs
 * per agent are counted in 32 bits); pt_batch_setup refuses larger batches. */
#define PT_BATCH_SLOT_TILE 8192
#define PT_BATCH_MAX_SLOTS 715827882

/* n replicates: params[r] and seed_keys[2r], seed_keys[2r+1] (the Threefry
 * key of replicate r, SEED[0..1] of setup_loadSeed PandemicSim.cu:2843-2862)
 * as pt_sim_create takes them. */
int pt_batch_create(int32_t n, const pt_params* params, const int32_t* seed_keys, int32_t device, pt_batch** out);

/* setupSim of every replicate (PandemicSim.cu:146-208, as pt_sim_setup).
 * n_people_out[r] = population of replicate r (may be NULL). */
int pt_batch_setup(pt_batch* b, int64_t* n_people_out);

/* setup_initialInfected for every replicate (PandemicSim.cu:708-758, as
 * pt_sim_seed_infections).  ids holds the replicates' id lists back to back:
 * replicate r contributes n_pandemic[r] + n_seasonal[r] ids, local to it. */
int pt_batch_seed_infections(pt_batch* b, const int32_t* ids, const int32_t* n_pandemic, const int32_t* n_seasonal);

/* Shape: out[0] replicates, out[1] D = the largest max_days (row stride of
 * pt_batch_run's outputs), out[2] agents of all replicates, out[3] agent
 * slots in HBM (each replicate padded to a multiple of 8192). */
int pt_batch_get_info(pt_batch* b, int64_t out[4]);

/* Runs every replicate's runToCompletion loop (PandemicSim.cu:877-967) to
 * its end.  counts_out[(r*D + day)*16 + k] and infected_out[r*D + day] as
 * pt_sim_run writes them per replicate (either may be NULL); days_out[r] =
 * rows of replicate r. */
int pt_batch_run(pt_batch* b, int32_t* counts_out, int32_t* infected_out, int32_t* days_out);

/* As pt_sim_generation_offsets, for replicate r. */
int pt_batch_generation_offsets(pt_batch* b, int32_t r, int32_t* c_pandemic, int32_t* c_seasonal, int32_t len);

/* Parity hooks: the per-replicate arrays of enum pt_array (PT_STATUS_P ..
 * PT_WP_PEOPLE and PT_HH_TYPES; the day tables are shared and not
 * exported). */
int64_t pt_batch_array_bytes(pt_batch* b, int32_t r, int32_t which);
int pt_batch_copy_array(pt_batch* b, int32_t r, int32_t which, void* dst, int64_t bytes);

/* Interventions of every replicate (iv[r], as pt_sim_set_interventions);
 * call before pt_batch_setup.  Default: none. */
int pt_batch_set_interventions(pt_batch* b, const pt_interventions* iv);

/* As pt_sim_set_table_mode; applies to every replicate. */
int pt_batch_set_table_mode(pt_batch* b, int32_t mode);

/* Day step as HIP graphs (default 1): pt_batch_run captures the weekday and
 * the weekend step once (again after setup or a table-mode change) and
 * replays one graph per day; 0 enqueues every kernel of every day instead.
 * Results are identical either way.  No counterpart in the reference,
 * which launches each day's kernels from the host (PandemicSim.cu:877-967). */
int pt_batch_use_graphs(pt_batch* b, int32_t enable);

/* Launch shapes (pt_launch_knob of pt_abi.h; PT_LAUNCH_GEN_BLOCKS_PER_CU is
 * unused here: the batch regenerates keys in its selection kernels).  A
 * change takes effect at the next pt_batch_run (graphs are recaptured). */
int pt_batch_set_launch(pt_batch* b, int32_t knob, int32_t value);
int pt_batch_get_launch(pt_batch* b, int32_t knob);

/* Device-side save / restore of every replicate's post-seeding state. */
int pt_batch_snapshot(pt_batch* b);
int pt_batch_restore(pt_batch* b);

int pt_batch_synchronize(pt_batch* b);
/* Loop time of the last pt_batch_run in milliseconds (HIP events). */
double pt_batch_last_run_ms(pt_batch* b);

void pt_batch_destroy(pt_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* PANDEMIC_PT_BATCH_H */
