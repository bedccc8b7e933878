// pt_batch.hip — replicate-batch engine (include/pandemic/pt_batch.h).
//
// Many independent simulations advance in lockstep: every day kernel covers
// all replicates at once, so a day costs one set of launches however many
// replicates there are.  This is the form the reference reaches only by
// running one process per constants row (PythonBatcher/).
//
// HBM layout.  Replicate r owns
//   agents      [abase_r, abase_r + N_r) of every per-agent array, abase_r a
//               multiple of the sweep tile (8192) so that a sweep block
//               never straddles two replicates;
//   households  hh_off[hbase_r .. hbase_r + H_r] (values local agent ids);
//   workplaces  wp_off[wbase_r .. wbase_r + W_r] (values index wp_people
//               from abase_r, which holds local ids);
//   buckets     [kb_r, kb_r + K_r) of the day's errand table, kb_r a
//               multiple of 32 so that its queried-bucket bits own whole
//               words.
// Everything a replicate computes uses its own local numbering through
// pointers offset by these bases (pt_day.h), so its arithmetic is the
// single-simulation engine's.  The day's infected list holds global agent
// ids grouped by replicate (ascending within each); inf_start[r] is where
// replicate r begins, and a thread finds its replicate by binary search.
// The errand tables of all replicates come from one sort of the selected
// (kb_r + key, local id) pairs; the per-replicate key ranges do not overlap,
// so each replicate's buckets are exactly its own table.
//
// Canonical infections: claim words are per agent and claim keys use the
// infector's position within its replicate's list, so the winner of every
// victim is the one the replicate's own run picks.  The claim lists are
// shared (global ids) and applied by one kernel.
//
// Citations are relative to PandemicThrust/PandemicThrust/ of the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pandemic/pt_abi.h"
#include "../../include/pandemic/pt_batch.h"
#include "pt_day.h"
#include "pt_model.h"
#include "pt_sort.h"

using namespace pt;

namespace {

// Where replicate r lives in the batch arrays.
struct RepInfo {
    int64_t abase;           // first agent slot
    int64_t hbase;           // first household-offset entry
    int64_t wbase;           // first workplace-offset entry
    int32_t N, H, W;
    int32_t tile0, ntiles;   // sweep tiles [tile0, tile0 + ntiles)
    uint32_t kb_wd, kb_we;   // first bucket in the weekday / weekend table
    int32_t pad;
    uint64_t workplace_off;  // workplace counter (H / 4)
};

// Base pointers of the batch arrays.
struct BatchArrays {
    const uint8_t* age;
    const int32_t* household;
    const int32_t* hh_off;
    const int32_t* wp_off;
    const int32_t* wp_people;
    const int32_t* er_off;
    const int32_t* er_people;
    const uint8_t *st_p, *st_s, *dy_p, *dy_s, *gn_p, *gn_s;
    uint64_t *claim_p, *claim_s;
};

// Replicate r's arrays in its local numbering.
__device__ __forceinline__ AgentArrays view(const BatchArrays& B, const RepInfo& ri, uint32_t kb) {
    AgentArrays A;
    A.age = B.age + ri.abase;
    A.household = B.household + ri.abase;
    A.hh_off = B.hh_off + ri.hbase;
    A.wp_off = B.wp_off + ri.wbase;
    A.wp_people = B.wp_people + ri.abase;
    A.er_off = B.er_off + kb;
    A.er_people = B.er_people;
    A.st_p = B.st_p + ri.abase;
    A.st_s = B.st_s + ri.abase;
    A.dy_p = B.dy_p + ri.abase;
    A.dy_s = B.dy_s + ri.abase;
    A.gn_p = B.gn_p + ri.abase;
    A.gn_s = B.gn_s + ri.abase;
    A.claim_p = B.claim_p + ri.abase;
    A.claim_s = B.claim_s + ri.abase;
    A.trace_victim = nullptr;
    A.trace_type = nullptr;
    A.trace_loc = nullptr;
    A.trace_attempt = nullptr;
    A.trace_draw = nullptr;
    A.trace_prof = nullptr;
    return A;
}

// Last r in [0, R) with start[r] <= pos: the replicate holding list
// position pos (replicates with no infected have start[r] == start[r+1]).
__device__ __forceinline__ int rep_of(const int32_t* __restrict__ start, int R, int64_t pos) {
    int lo = 0, hi = R;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if ((int64_t)start[mid] <= pos)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

constexpr int kRepThreads = 256;

// D1 + D2 of every running replicate.
__global__ __launch_bounds__(kSweepThreads) void kb_sweep(const DayState* __restrict__ g, const DayState* __restrict__ sts,
                                                          const RepInfo* __restrict__ reps, const Model* __restrict__ models,
                                                          const int32_t* __restrict__ tile_rep, uint8_t* __restrict__ st_p,
                                                          uint8_t* __restrict__ st_s, const uint8_t* __restrict__ dy_p,
                                                          const uint8_t* __restrict__ dy_s,
                                                          uint32_t* __restrict__ blk_counts,
                                                          uint32_t* __restrict__ blk_ninf, int32_t* __restrict__ staging) {
    if (g->done) return;
    const int64_t b = blockIdx.x;
    const int r = tile_rep[b];
    if (sts[r].done) return;
    const RepInfo ri = reps[r];
    const int64_t a = ri.abase;
    sweep_block(models[r], ri.N, sts[r].day, (b - ri.tile0) * (int64_t)kSweepTile, st_p + a, st_s + a, dy_p + a,
                dy_s + a,
                blk_counts + b * 16, blk_ninf + b, staging + b * kSweepTile, (int32_t)a);
}

// Day plan of replicate blockIdx.x (the single-simulation k_plan): daily
// counts, infected count and local block offsets, the stop rule
// (PandemicSim.cu:899-901), the table path and the day's RNG offsets
// (:3030-3040, :1409-1415, :1278-1279, :1393-1394).
__global__ __launch_bounds__(kRepThreads) void kb_plan(const DayState* __restrict__ g, DayState* __restrict__ sts,
                                                       const RepInfo* __restrict__ reps, const Model* __restrict__ models,
                                                       const uint32_t* __restrict__ blk_counts,
                                                       const uint32_t* __restrict__ blk_ninf,
                                                       uint32_t* __restrict__ blk_off, int32_t* __restrict__ rep_I,
                                                       int32_t* __restrict__ out_counts, int32_t* __restrict__ out_inf,
                                                       int D, int table_mode, int mark_pct) {
    if (g->done) return;
    const int r = blockIdx.x;
    DayState* st = sts + r;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (st->done) {
        if (tid == 0) rep_I[r] = 0;
        return;
    }
    constexpr int NW = kRepThreads / 64;
    __shared__ uint32_t s_red[NW][16];
    __shared__ uint32_t s_w[NW];
    __shared__ uint32_t s_carry;
    const RepInfo ri = reps[r];
    const int t0 = ri.tile0, nt = ri.ntiles;
    uint32_t c[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) c[i] = 0;
    for (int b = tid; b < nt; b += kRepThreads)
#pragma unroll
        for (int i = 0; i < 16; ++i) c[i] += blk_counts[(int64_t)(t0 + b) * 16 + i];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const uint32_t v = wave_sum(c[i]);
        if (lane == 0) s_red[wave][i] = v;
    }
    if (tid == 0) s_carry = 0;
    __syncthreads();
    const int day = st->day;
    if (tid < 16) {
        uint32_t s = 0;
        for (int w = 0; w < NW; ++w) s += s_red[w][tid];
        out_counts[((int64_t)r * D + day) * 16 + tid] = (int32_t)s;
    }
    for (int b0 = 0; b0 < nt; b0 += kRepThreads) {
        const int b = b0 + tid;
        const uint32_t v = b < nt ? blk_ninf[t0 + b] : 0u;
        const uint32_t inc = wave_incl_scan(v, lane);
        if (lane == 63) s_w[wave] = inc;
        __syncthreads();
        uint32_t wbase = 0, tot = 0;
        for (int w = 0; w < NW; ++w) {
            if (w < wave) wbase += s_w[w];
            tot += s_w[w];
        }
        const uint32_t carry = s_carry;
        if (b < nt) blk_off[t0 + b] = carry + wbase + inc - v;
        __syncthreads();
        if (tid == 0) s_carry = carry + tot;
        __syncthreads();
    }
    if (tid == 0) {
        const int32_t I = (int32_t)s_carry;
        st->n_infected = I;
        st->days_run = day + 1;
        out_inf[(int64_t)r * D + day] = I;
        if (I == 0) {
            st->done = 1;
            st->skip_mark = 1;
            rep_I[r] = 0;
        } else {
            const Model& M = models[r];
            const bool weekend = (day % 7) >= 5;
            const int32_t K = weekend ? kWeHours * M.we_E : M.wd_K;
            const bool mark = table_mode == TABLES_SUBSET || (table_mode == TABLES_AUTO && I <= mark_limit(K, mark_pct));
            st->skip_mark = mark ? 0 : 1;
            st->skip_full = mark ? 1 : 0;
            const int64_t N = ri.N;
            st->errand_off = st->rand_offset;
            st->rand_offset += weekend ? (uint64_t)(2 * N) : (uint64_t)(N / 2);
            st->contact_base = st->rand_offset;
            st->rand_offset += (uint64_t)6 * (uint64_t)I;
            rep_I[r] = I;
        }
    }
}

// Where each replicate's infected begin in the day's list; ends the batch
// when no replicate is running.
__global__ __launch_bounds__(1024) void kb_start(DayState* __restrict__ g, const int32_t* __restrict__ rep_I,
                                                 int32_t* __restrict__ inf_start, int R) {
    if (g->done) return;
    __shared__ uint32_t s_w[16];
    __shared__ uint32_t s_carry;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_carry = 0;
    __syncthreads();
    for (int r0 = 0; r0 < R; r0 += 1024) {
        const int r = r0 + tid;
        const uint32_t v = r < R ? (uint32_t)rep_I[r] : 0u;
        const uint32_t inc = wave_incl_scan(v, lane);
        if (lane == 63) s_w[wave] = inc;
        __syncthreads();
        uint32_t wbase = 0, tot = 0;
        for (int w = 0; w < 16; ++w) {
            if (w < wave) wbase += s_w[w];
            tot += s_w[w];
        }
        const uint32_t carry = s_carry;
        if (r < R) inf_start[r] = (int32_t)(carry + wbase + inc - v);
        __syncthreads();
        if (tid == 0) s_carry = carry + tot;
        __syncthreads();
    }
    if (tid == 0) {
        const uint32_t total = s_carry;
        inf_start[R] = (int32_t)total;
        g->n_infected = (int32_t)total;
        if (total == 0) g->done = 1;  // a running replicate always has I > 0 after its plan
    }
}

__global__ void kb_compact(const DayState* __restrict__ g, const DayState* __restrict__ sts,
                           const int32_t* __restrict__ tile_rep, const int32_t* __restrict__ staging,
                           const uint32_t* __restrict__ blk_ninf, const uint32_t* __restrict__ blk_off,
                           const int32_t* __restrict__ inf_start, int32_t* __restrict__ infected) {
    if (g->done) return;
    const int64_t b = blockIdx.x;
    const int r = tile_rep[b];
    if (sts[r].done) return;
    const uint32_t n = blk_ninf[b];
    int32_t* dst = infected + inf_start[r] + blk_off[b];
    const int32_t* src = staging + b * kSweepTile;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

__global__ void kb_clear_marks(const DayState* __restrict__ g, uint32_t* __restrict__ mark, int64_t nwords) {
    if (g->done) return;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += stride) mark[i] = 0u;
}

template <bool WEEKEND>
__global__ __launch_bounds__(256) void kb_mark(const DayState* __restrict__ g, const DayState* __restrict__ sts,
                                               const RepInfo* __restrict__ reps, const Model* __restrict__ models,
                                               const int32_t* __restrict__ inf_start, int R,
                                               const int32_t* __restrict__ infected, BatchArrays B,
                                               uint32_t* __restrict__ mark) {
    if (g->done) return;
    const int64_t I = g->n_infected;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; pos < I; pos += stride) {
        const int r = rep_of(inf_start, R, pos);
        const DayState& st = sts[r];
        if (st.skip_mark) continue;
        const RepInfo ri = reps[r];
        const uint32_t kb = WEEKEND ? ri.kb_we : ri.kb_wd;
        const AgentArrays A = view(B, ri, kb);
        const int32_t idx = infected[pos] - (int32_t)ri.abase;
        if (WEEKEND)
            mark_weekend(models[r], A, st.day, st.contact_base, st.errand_off, pos - inf_start[r], idx, mark, kb);
        else
            mark_weekday(models[r], A, st.day, st.errand_off, idx, mark, kb);
    }
}

// Table path per replicate: every pair (skip_full = 0) or the pairs of its
// queried buckets (skip_full = 1), as the single-simulation k_decide.
__global__ __launch_bounds__(kRepThreads) void kb_decide(const DayState* __restrict__ g, DayState* __restrict__ sts,
                                                         const RepInfo* __restrict__ reps,
                                                         const Model* __restrict__ models,
                                                         const uint32_t* __restrict__ mark, int weekend, int table_mode,
                                                         int share_pct) {
    if (g->done) return;
    const int r = blockIdx.x;
    DayState* st = sts + r;
    if (st->done || st->skip_mark) return;
    const Model& M = models[r];
    const int64_t K = weekend ? (int64_t)kWeHours * M.we_E : (int64_t)M.wd_K;
    const uint32_t kb = weekend ? reps[r].kb_we : reps[r].kb_wd;
    const uint32_t* w = mark + kb / 32;
    const int64_t nwords = (K + 31) / 32;
    uint32_t c = 0;
    for (int64_t i = threadIdx.x; i < nwords; i += kRepThreads) c += (uint32_t)__popc(w[i]);
    c = wave_sum(c);
    __shared__ uint32_t s_w[kRepThreads / 64];
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m = 0;
        for (int i = 0; i < kRepThreads / 64; ++i) m += s_w[i];
        st->n_marked = m;
        const bool subset = table_mode == TABLES_SUBSET || (table_mode == TABLES_AUTO && subset_pays(m, K, share_pct));
        st->skip_full = subset ? 1 : 0;
    }
}

// Selection, first pass.  A block of a replicate that takes every pair today
// (skip_full == 0) only records its pair count; a block of a replicate that
// takes its queried buckets stages those pairs (select_block) and records
// their count.
template <int PER, bool WEEKEND>
__global__ __launch_bounds__(kSelThreads) void kb_select(const DayState* __restrict__ g, DayState* __restrict__ sts,
                                                         const RepInfo* __restrict__ reps,
                                                         const Model* __restrict__ models,
                                                         const int32_t* __restrict__ tile_rep,
                                                         const uint8_t* __restrict__ age, const uint32_t* __restrict__ mark,
                                                         uint32_t* __restrict__ stage_keys, int32_t* __restrict__ stage_ids,
                                                         uint32_t* __restrict__ blk_cnt) {
    if (g->done) return;
    constexpr int KPA = WEEKEND ? 3 : 2;
    constexpr int AG = kSelThreads * PER;
    constexpr int STAGE = AG * KPA;
    const int64_t b = blockIdx.x;
    const int64_t a = b * AG;
    const int r = tile_rep[a / kSweepTile];
    DayState* st = sts + r;
    const RepInfo ri = reps[r];
    const int64_t la = a - ri.abase;
    if (st->done || la >= ri.N) {
        if (threadIdx.x == 0) blk_cnt[b] = 0;
        return;
    }
    if (st->skip_full == 0) {
        if (threadIdx.x == 0) blk_cnt[b] = (uint32_t)(KPA * min((int64_t)AG, (int64_t)ri.N - la));
        return;
    }
    if (!select_block<PER, WEEKEND>(models[r], st->errand_off, age + ri.abase, mark, WEEKEND ? ri.kb_we : ri.kb_wd,
                                    false, la, ri.N, stage_keys + b * STAGE, stage_ids + b * STAGE, blk_cnt + b))
        st->fault = 1;
}

// Selection, second pass: every block writes its pairs at its offset in the
// day's pair list — regenerated for every-pair replicates, copied from the
// staging area for queried-bucket replicates.
template <int PER, bool WEEKEND>
__global__ __launch_bounds__(kSelThreads) void kb_emit(const DayState* __restrict__ g, DayState* __restrict__ sts,
                                                       const RepInfo* __restrict__ reps, const Model* __restrict__ models,
                                                       const int32_t* __restrict__ tile_rep,
                                                       const uint8_t* __restrict__ age, const uint32_t* __restrict__ mark,
                                                       const uint32_t* __restrict__ stage_keys,
                                                       const int32_t* __restrict__ stage_ids,
                                                       const uint32_t* __restrict__ blk_cnt,
                                                       const uint32_t* __restrict__ blk_off, uint32_t* __restrict__ sel_keys,
                                                       int32_t* __restrict__ sel_ids, const uint32_t* __restrict__ seg_start) {
    if (g->done) return;
    constexpr int KPA = WEEKEND ? 3 : 2;
    constexpr int AG = kSelThreads * PER;
    constexpr int STAGE = AG * KPA;
    const int64_t b = blockIdx.x;
    const int64_t a = b * AG;
    const int r = tile_rep[a / kSweepTile];
    DayState* st = sts + r;
    const RepInfo ri = reps[r];
    const int64_t la = a - ri.abase;
    if (st->done || la >= ri.N) return;
    // Global keys and dense positions, or (segmented tables) keys local to
    // the replicate at its segment's positions.
    const uint32_t kb = WEEKEND ? ri.kb_we : ri.kb_wd;
    uint32_t off = blk_off[b], kbias = kb, ksub = 0;
    if (seg_start) {
        off = seg_start[r] + (off - blk_off[ri.abase / AG]);
        kbias = 0;
        ksub = kb;
    }
    if (st->skip_full == 0) {
        if (!select_block<PER, WEEKEND>(models[r], st->errand_off, age + ri.abase, mark, kbias, true, la, ri.N,
                                        sel_keys + off, sel_ids + off, nullptr))
            st->fault = 1;
        return;
    }
    const uint32_t n = blk_cnt[b];
    for (uint32_t i = threadIdx.x; i < n; i += kSelThreads) {
        sel_keys[off + i] = stage_keys[b * STAGE + i] - ksub;
        sel_ids[off + i] = stage_ids[b * STAGE + i];
    }
}

__global__ void kb_set_nsel(DayState* __restrict__ g, const uint32_t* __restrict__ blk_cnt,
                            const uint32_t* __restrict__ blk_off, int nblk) {
    if (g->done) return;
    g->n_sel = blk_off[nblk - 1] + blk_cnt[nblk - 1];
}

// Segmented tables: each replicate's pair range for the day (the pairs of
// its selection blocks, AG agents each, moved to a start rounded up to a
// multiple of 4) and the day's pair count including those gaps.  One block
// of kSegBlocks threads (R <= kSegBlocks).
static_assert(kSweepTile % kSelAgentsWd == 0 && kSweepTile % kSelAgentsWe == 0,
              "replicates (whole sweep tiles) start on selection-block boundaries");
template <int AG>
__global__ __launch_bounds__(kSegBlocks) void kb_segments(DayState* __restrict__ g, const RepInfo* __restrict__ reps,
                                                          int R, int nb, const uint32_t* __restrict__ blk_cnt,
                                                          const uint32_t* __restrict__ blk_off,
                                                          uint32_t* __restrict__ seg_start,
                                                          uint32_t* __restrict__ seg_count) {
    if (g->done) return;
    constexpr int NW = kSegBlocks / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ uint32_t s_w[NW];
    uint32_t cnt = 0;
    if (tid < R) {
        const int64_t b0 = reps[tid].abase / AG;
        const int64_t b1 = tid + 1 < R ? reps[tid + 1].abase / AG : (int64_t)nb;
        if (b1 > b0) cnt = blk_off[b1 - 1] + blk_cnt[b1 - 1] - blk_off[b0];
    }
    const uint32_t al = (cnt + 3u) & ~3u;
    const uint32_t inc = wave_incl_scan(al, lane);
    if (lane == 63) s_w[wave] = inc;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        if (w < wave) base += s_w[w];
        tot += s_w[w];
    }
    if (tid < R) {
        seg_start[tid] = base + inc - al;
        seg_count[tid] = cnt;
    }
    if (tid == 0) g->n_sel = tot;
}

template <bool WEEKEND>
__global__ __launch_bounds__(kContactThreads) void kb_contacts(DayState* __restrict__ g, const DayState* __restrict__ sts,
                                                               const RepInfo* __restrict__ reps,
                                                               const Model* __restrict__ models,
                                                               const int32_t* __restrict__ inf_start, int R,
                                                               const int32_t* __restrict__ infected, BatchArrays B,
                                                               int32_t* __restrict__ list_p, int32_t* __restrict__ list_s) {
    if (g->done) return;
    const int64_t I = g->n_infected;
    const int lane = threadIdx.x & 63;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t b0 = (int64_t)blockIdx.x * blockDim.x; b0 < I; b0 += stride) {
        const int64_t pos = b0 + threadIdx.x;
        int32_t vic[8];
        int typ[8];
        bool newp[8], news[8];
        clear_contacts(vic, typ, newp, news);
        int32_t bias = 0;
        if (pos < I) {
            const int r = rep_of(inf_start, R, pos);
            const DayState& st = sts[r];
            const RepInfo ri = reps[r];
            const AgentArrays A = view(B, ri, WEEKEND ? ri.kb_we : ri.kb_wd);
            const int32_t idx = infected[pos] - (int32_t)ri.abase;
            const int64_t lpos = pos - inf_start[r];
            bias = (int32_t)ri.abase;
            if (WEEKEND)
                weekend_contacts(models[r], A, st.day, st.contact_base, st.errand_off, idx, lpos, vic, typ, newp, news);
            else
                weekday_contacts(models[r], A, st.day, st.contact_base, st.errand_off, ri.workplace_off, idx, lpos, vic,
                                 typ, newp, news);
        }
        append_claims(g->n_claimed, list_p, list_s, bias, vic, newp, news, lane);
    }
}

__global__ void kb_end_day(DayState* __restrict__ g, DayState* __restrict__ sts, int R) {
    if (g->done) return;
    for (int r = threadIdx.x; r < R; r += blockDim.x) {
        DayState* st = sts + r;
        if (st->done) continue;
        st->day += 1;
        if (st->day >= st->max_days) st->done = 1;
    }
    if (threadIdx.x == 0) {
        g->day += 1;
        g->n_claimed[0] = 0;
        g->n_claimed[1] = 0;
    }
}

__global__ void kb_fill_u64(uint64_t* p, int64_t n, uint64_t v) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

}  // namespace

struct pt_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int R = 0;
    int D = 0;  // largest max_days
    int table_mode = TABLES_AUTO;
    std::vector<pt_params> params;
    std::vector<Model> models;
    std::vector<RepInfo> reps;
    std::vector<uint64_t> rand_after;  // rand_offset after setup, per replicate
    int64_t P = 0, HH = 0, WW = 0, n_agents = 0;
    int ntiles = 0;
    uint32_t K_wd = 0, K_we = 0;  // buckets of the batch tables
    bool setup_done = false, have_snapshot = false;
    // device
    Model* d_models = nullptr;
    RepInfo* d_reps = nullptr;
    DayState* d_states = nullptr;
    DayState* d_g = nullptr;
    int32_t* tile_rep = nullptr;
    uint8_t *st_p = nullptr, *st_s = nullptr, *dy_p = nullptr, *dy_s = nullptr, *gn_p = nullptr, *gn_s = nullptr;
    uint8_t *age = nullptr, *hh_types = nullptr, *snap = nullptr;
    int32_t *household = nullptr, *hh_off = nullptr, *wp_off = nullptr, *wp_people = nullptr;
    uint64_t *claim_p = nullptr, *claim_s = nullptr;
    int32_t *list_p = nullptr, *list_s = nullptr, *infected = nullptr, *staging = nullptr;
    uint32_t *blk_counts = nullptr, *blk_ninf = nullptr, *blk_off = nullptr;
    int32_t *rep_I = nullptr, *inf_start = nullptr;
    uint32_t* mark = nullptr;
    uint32_t *stage_keys = nullptr, *sel_keys = nullptr;
    int32_t *stage_ids = nullptr, *sel_ids = nullptr;
    uint32_t *sblk_cnt = nullptr, *sblk_off = nullptr;
    int32_t *er_people = nullptr, *er_off = nullptr;
    // Segmented day tables (R <= kSegBlocks): each replicate's pairs sorted
    // in place over its local keys.  Static: the radix blocks of each
    // replicate and its bucket ranges; per day: its pair ranges.
    bool seg_tables = false;
    int32_t *seg_blk = nullptr, *seg_blk0 = nullptr, *seg_nblk = nullptr;
    uint32_t *seg_kb_wd = nullptr, *seg_kb_we = nullptr;  // R + 1
    int32_t *seg_K_wd = nullptr, *seg_K_we = nullptr;     // R
    uint32_t *seg_start = nullptr, *seg_count = nullptr;  // R
    int32_t K_local_wd = 0, K_local_we = 0;
    int32_t *d_counts = nullptr, *d_inf = nullptr;
    unsigned long long* d_scratch64 = nullptr;  // max(R, 512)
    TableScratch tables;
    int grid_contacts = 2048;
    int cus = 256;                                   // compute units of the device
    int launch[PT_LAUNCH_KNOBS] = {8, 16, 8, 100, 75, 9, 1};  // pt_launch_knob values
    hipEvent_t run_begin = nullptr, run_end = nullptr;
    // pt_batch_run's early stop: the batch's `done` flag copied back after
    // each group of days (pinned, two slots) and the groups' events.
    int32_t* h_done = nullptr;
    hipEvent_t group_ev[2] = {nullptr, nullptr};
    // The day step captured once per layout as HIP graphs (weekday,
    // weekend) and replayed for every day; the kernels take the day and
    // every count from the device, so one capture serves all days.
    bool use_graphs = true;
    hipGraphExec_t day_graph[2] = {nullptr, nullptr};
    int graph_table_mode = -1;
    double last_run_ms = 0;
};

namespace {

void free_batch_population(pt_batch* b) {
    dev_free(b->tile_rep);
    dev_free(b->st_p);
    dev_free(b->st_s);
    dev_free(b->dy_p);
    dev_free(b->dy_s);
    dev_free(b->gn_p);
    dev_free(b->gn_s);
    dev_free(b->age);
    dev_free(b->hh_types);
    dev_free(b->snap);
    dev_free(b->household);
    dev_free(b->hh_off);
    dev_free(b->wp_off);
    dev_free(b->wp_people);
    dev_free(b->claim_p);
    dev_free(b->claim_s);
    dev_free(b->list_p);
    dev_free(b->list_s);
    dev_free(b->infected);
    dev_free(b->staging);
    dev_free(b->blk_counts);
    dev_free(b->blk_ninf);
    dev_free(b->blk_off);
    dev_free(b->mark);
    dev_free(b->stage_keys);
    dev_free(b->stage_ids);
    dev_free(b->sel_keys);
    dev_free(b->sel_ids);
    dev_free(b->sblk_cnt);
    dev_free(b->sblk_off);
    dev_free(b->er_people);
    dev_free(b->er_off);
    dev_free(b->seg_blk);
    dev_free(b->seg_blk0);
    dev_free(b->seg_nblk);
    dev_free(b->seg_kb_wd);
    dev_free(b->seg_kb_we);
    dev_free(b->seg_K_wd);
    dev_free(b->seg_K_we);
    dev_free(b->seg_start);
    dev_free(b->seg_count);
    b->seg_tables = false;
    dev_free(b->d_counts);
    dev_free(b->d_inf);
    table_scratch_free(b->tables);
    b->setup_done = b->have_snapshot = false;
}

int reset_batch_state(pt_batch* b) {
    std::vector<DayState> h(b->R);
    for (int r = 0; r < b->R; ++r) {
        DayState& s = h[r];
        memset(&s, 0, sizeof s);
        s.done = b->params[r].max_days <= 0 ? 1 : 0;
        s.rand_offset = b->rand_after[r];
        s.max_days = b->params[r].max_days;
        s.skip_mark = 1;
        s.skip_full = 1;
    }
    DayState g;
    memset(&g, 0, sizeof g);
    HIP_TRY(hipMemcpyAsync(b->d_states, h.data(), sizeof(DayState) * b->R, hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->d_g, &g, sizeof g, hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

BatchArrays batch_arrays(pt_batch* b) {
    BatchArrays B;
    B.age = b->age;
    B.household = b->household;
    B.hh_off = b->hh_off;
    B.wp_off = b->wp_off;
    B.wp_people = b->wp_people;
    B.er_off = b->er_off;
    B.er_people = b->er_people;
    B.st_p = b->st_p;
    B.st_s = b->st_s;
    B.dy_p = b->dy_p;
    B.dy_s = b->dy_s;
    B.gn_p = b->gn_p;
    B.gn_s = b->gn_s;
    B.claim_p = b->claim_p;
    B.claim_s = b->claim_s;
    return B;
}

// One lockstep day d of every replicate.
int enqueue_batch_day(pt_batch* b, int d) {
    const bool weekend = (d % 7) >= 5;
    hipStream_t s = b->stream;
    DayState* g = b->d_g;
    const int R = b->R;
    kb_sweep<<<b->ntiles, kSweepThreads, 0, s>>>(g, b->d_states, b->d_reps, b->d_models, b->tile_rep, b->st_p, b->st_s, b->dy_p,
                                                 b->dy_s, b->blk_counts, b->blk_ninf, b->staging);
    kb_plan<<<R, kRepThreads, 0, s>>>(g, b->d_states, b->d_reps, b->d_models, b->blk_counts, b->blk_ninf, b->blk_off,
                                      b->rep_I, b->d_counts, b->d_inf, b->D, b->table_mode,
                                      b->launch[PT_LAUNCH_MARK_LIMIT]);
    kb_start<<<1, 1024, 0, s>>>(g, b->rep_I, b->inf_start, R);
    kb_compact<<<b->ntiles, 256, 0, s>>>(g, b->d_states, b->tile_rep, b->staging, b->blk_ninf, b->blk_off,
                                         b->inf_start, b->infected);
    const uint32_t K = weekend ? b->K_we : b->K_wd;
    const int64_t nwords = ((int64_t)K + 31) / 32;
    const BatchArrays B = batch_arrays(b);
    kb_clear_marks<<<grid_for(nwords, 256, 1024), 256, 0, s>>>(g, b->mark, nwords);
    if (weekend)
        kb_mark<true><<<b->grid_contacts, 256, 0, s>>>(g, b->d_states, b->d_reps, b->d_models, b->inf_start, R,
                                                       b->infected, B, b->mark);
    else
        kb_mark<false><<<b->grid_contacts, 256, 0, s>>>(g, b->d_states, b->d_reps, b->d_models, b->inf_start, R,
                                                        b->infected, B, b->mark);
    kb_decide<<<R, kRepThreads, 0, s>>>(g, b->d_states, b->d_reps, b->d_models, b->mark, weekend ? 1 : 0,
                                        b->table_mode, b->launch[PT_LAUNCH_SUBSET_SHARE]);
    const int nb = (int)(b->P / (weekend ? kSelAgentsWe : kSelAgentsWd));
    if (weekend)
        kb_select<kSelPerWe, true><<<nb, kSelThreads, 0, s>>>(g, b->d_states, b->d_reps, b->d_models, b->tile_rep,
                                                               b->age, b->mark, b->stage_keys, b->stage_ids,
                                                               b->sblk_cnt);
    else
        kb_select<kSelPerWd, false><<<nb, kSelThreads, 0, s>>>(g, b->d_states, b->d_reps, b->d_models, b->tile_rep,
                                                                b->age, b->mark, b->stage_keys, b->stage_ids,
                                                                b->sblk_cnt);
    hipError_t e = exclusive_scan_u32(b->sblk_cnt, b->sblk_off, nb, b->tables.scan_part, &g->done, s);
    if (e != hipSuccess) return fail(PT_EDEVICE, "selection scan: %s", hipGetErrorString(e));
    const bool seg = b->seg_tables && b->launch[PT_LAUNCH_SEGMENTED_TABLES] != 0;
    const uint32_t* seg_start = seg ? b->seg_start : nullptr;
    if (!seg)
        kb_set_nsel<<<1, 1, 0, s>>>(g, b->sblk_cnt, b->sblk_off, nb);
    else if (weekend)
        kb_segments<kSelAgentsWe><<<1, kSegBlocks, 0, s>>>(g, b->d_reps, R, nb, b->sblk_cnt, b->sblk_off, b->seg_start,
                                                           b->seg_count);
    else
        kb_segments<kSelAgentsWd><<<1, kSegBlocks, 0, s>>>(g, b->d_reps, R, nb, b->sblk_cnt, b->sblk_off, b->seg_start,
                                                           b->seg_count);
    if (weekend)
        kb_emit<kSelPerWe, true><<<nb, kSelThreads, 0, s>>>(g, b->d_states, b->d_reps, b->d_models, b->tile_rep, b->age,
                                                             b->mark, b->stage_keys, b->stage_ids, b->sblk_cnt, b->sblk_off,
                                                             b->sel_keys, b->sel_ids, seg_start);
    else
        kb_emit<kSelPerWd, false><<<nb, kSelThreads, 0, s>>>(g, b->d_states, b->d_reps, b->d_models, b->tile_rep,
                                                              b->age, b->mark, b->stage_keys, b->stage_ids, b->sblk_cnt,
                                                              b->sblk_off, b->sel_keys, b->sel_ids, seg_start);
    const int64_t n_cap = (weekend ? 3 : 2) * b->P + 4 * (int64_t)R;
    if (seg) {
        SegLayout sl;
        sl.blk_seg = b->seg_blk;
        sl.seg_blk0 = b->seg_blk0;
        sl.seg_nblk = b->seg_nblk;
        sl.seg_start = b->seg_start;
        sl.seg_count = b->seg_count;
        sl.seg_kbase = weekend ? b->seg_kb_we : b->seg_kb_wd;
        sl.seg_K = weekend ? b->seg_K_we : b->seg_K_wd;
        sl.R = R;
        e = build_table_segmented(b->tables, b->sel_keys, b->sel_ids, n_cap, weekend ? b->K_local_we : b->K_local_wd,
                                  (int32_t)K, sl, b->er_people, b->er_off, &g->done, s, nullptr);
    } else {
        e = build_table(b->tables, b->sel_keys, b->sel_ids, n_cap, &g->n_sel, (int32_t)K, ID_IDENTITY, 0,
                        b->er_people, b->er_off, &g->done, s, nullptr);
    }
    if (e != hipSuccess) return fail(PT_EDEVICE, "batch table build: %s", hipGetErrorString(e));
    if (weekend)
        kb_contacts<true><<<b->grid_contacts, kContactThreads, 0, s>>>(g, b->d_states, b->d_reps, b->d_models,
                                                                       b->inf_start, R, b->infected, B, b->list_p,
                                                                       b->list_s);
    else
        kb_contacts<false><<<b->grid_contacts, kContactThreads, 0, s>>>(g, b->d_states, b->d_reps, b->d_models,
                                                                        b->inf_start, R, b->infected, B, b->list_p,
                                                                        b->list_s);
    k_apply<<<b->grid_contacts, 256, 0, s>>>(g, b->claim_p, b->claim_s, b->list_p, b->list_s, b->st_p, b->st_s,
                                             b->dy_p, b->dy_s, b->gn_p, b->gn_s);
    kb_end_day<<<1, 256, 0, s>>>(g, b->d_states, R);
    e = hipGetLastError();
    if (e != hipSuccess) return fail(PT_EDEVICE, "batch day launch: %s", hipGetErrorString(e));
    return 0;
}

void drop_day_graphs(pt_batch* b) {
    for (auto& g : b->day_graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    b->graph_table_mode = -1;
}

// Captures the weekday (day 0) and weekend (day 5) steps.
int capture_day_graphs(pt_batch* b) {
    drop_day_graphs(b);
    for (int w = 0; w < 2; ++w) {
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_batch_day(b, w ? 5 : 0);
        const hipError_t e = hipStreamEndCapture(b->stream, &graph);
        if (rc != 0 || e != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            drop_day_graphs(b);
            return rc != 0 ? rc : fail(PT_EDEVICE, "day capture: %s", hipGetErrorString(e));
        }
        const hipError_t ei = hipGraphInstantiate(&b->day_graph[w], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            drop_day_graphs(b);
            return fail(PT_EDEVICE, "day graph: %s", hipGetErrorString(ei));
        }
    }
    b->graph_table_mode = b->table_mode;
    return 0;
}

}  // namespace

// ============================ C ABI ============================
extern "C" {

int pt_batch_create(int32_t n, const pt_params* params, const int32_t* seed_keys, int32_t device, pt_batch** out) {
    if (!params || !seed_keys || !out) return fail(PT_EINVAL, "null argument");
    *out = nullptr;
    if (n <= 0 || n > 65536) return fail(PT_EINVAL, "replicate count %d outside [1, 65536]", n);
    std::vector<Model> models(n);
    int D = 0;
    for (int r = 0; r < n; ++r) {
        const int rc = build_model(&params[r], seed_keys + 2 * r, models[r]);
        if (rc != PT_OK) return fail(rc, "replicate %d: %s", r, pt_last_error());
        D = std::max(D, params[r].max_days);
    }
    HIP_TRY(hipSetDevice(device));
    pt_batch* b = new pt_batch();
    b->device = device;
    b->R = n;
    b->D = std::max(D, 1);
    b->params.assign(params, params + n);
    b->models = models;
    b->reps.resize(n);
    b->rand_after.assign(n, 0);
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete b;
        return fail(PT_EDEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        b->cus = cus;
    b->grid_contacts = b->cus * b->launch[PT_LAUNCH_CONTACT_BLOCKS_PER_CU];
    if (dev_alloc(&b->d_models, n) || dev_alloc(&b->d_reps, n) || dev_alloc(&b->d_states, n) || dev_alloc(&b->d_g, 1) ||
        dev_alloc(&b->rep_I, n) || dev_alloc(&b->inf_start, (int64_t)n + 1) ||
        dev_alloc(&b->d_scratch64, std::max(n, 512))) {
        const std::string msg = pt_last_error();
        pt_batch_destroy(b);
        return fail(PT_ENOMEM, "%s", msg.c_str());
    }
    for (auto& e : b->group_ev) (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
    if (hipHostMalloc((void**)&b->h_done, 2 * sizeof(int32_t), hipHostMallocDefault) != hipSuccess) {
        b->h_done = nullptr;
        pt_batch_destroy(b);
        return fail(PT_ENOMEM, "pinned host allocation failed");
    }
    (void)hipEventCreate(&b->run_begin);
    (void)hipEventCreate(&b->run_end);
    *out = b;
    return PT_OK;
}

int pt_batch_setup(pt_batch* b, int64_t* n_people_out) {
    if (!b) return fail(PT_EINVAL, "null handle");
    HIP_TRY(hipSetDevice(b->device));
    drop_day_graphs(b);  // captured with the buffers freed below
    free_batch_population(b);
    const int R = b->R;
    hipStream_t st = b->stream;
    // K1/K2 of every replicate: household types and population sizes.
    int64_t HH = 0, WW = 0;
    for (int r = 0; r < R; ++r) {
        b->reps[r].hbase = HH;
        b->reps[r].wbase = WW;
        b->reps[r].H = b->models[r].H;
        b->reps[r].W = b->models[r].W;
        HH += (int64_t)b->models[r].H + 1;
        WW += (int64_t)b->models[r].W + 1;
        b->models[r].N = 0;
    }
    b->HH = HH;
    b->WW = WW;
    if (dev_alloc(&b->hh_types, HH)) return PT_ENOMEM;
    HIP_TRY(hipMemcpyAsync(b->d_models, b->models.data(), sizeof(Model) * R, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemsetAsync(b->d_scratch64, 0, sizeof(unsigned long long) * R, st));
    for (int r = 0; r < R; ++r)
        launch_household_types(b->d_models + r, b->reps[r].H, b->hh_types + b->reps[r].hbase, b->d_scratch64 + r, st);
    std::vector<unsigned long long> people(R);
    HIP_TRY(hipMemcpyAsync(people.data(), b->d_scratch64, sizeof(unsigned long long) * R, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    // Layout.
    int64_t P = 0, n_agents = 0, Nmax = 0, Hmax = 0;
    int tiles = 0;
    uint64_t kwd = 0, kwe = 0;
    for (int r = 0; r < R; ++r) {
        const unsigned long long np = people[r];
        if (np == 0 || np > (unsigned long long)(0x7FFFFFFF / 8))
            return fail(PT_EINVAL, "replicate %d: population %llu outside the supported range", r, np);
        const int64_t N = (int64_t)np;
        if ((int64_t)b->params[r].init_pandemic + b->params[r].init_seasonal > N)
            return fail(PT_EINVAL, "replicate %d: initial infected exceed the population", r);
        RepInfo& ri = b->reps[r];
        Model& m = b->models[r];
        m.N = (int32_t)N;
        ri.N = (int32_t)N;
        ri.abase = P;
        ri.tile0 = tiles;
        ri.ntiles = (int32_t)((N + kSweepTile - 1) / kSweepTile);
        tiles += ri.ntiles;
        P += (int64_t)ri.ntiles * kSweepTile;
        n_agents += N;
        Nmax = std::max(Nmax, N);
        Hmax = std::max<int64_t>(Hmax, m.H);
        ri.kb_wd = (uint32_t)kwd;
        ri.kb_we = (uint32_t)kwe;
        // At least one unqueried bucket after each replicate's: in segmented
        // tables it takes the alignment gap after the replicate's pairs.
        kwd += ((uint64_t)m.wd_K + 32) / 32 * 32;
        kwe += ((uint64_t)kWeHours * m.we_E + 32) / 32 * 32;
        ri.workplace_off = (uint64_t)(m.H / 4);
        b->rand_after[r] = ri.workplace_off + (uint64_t)(N / 4);
    }
    // Global agent ids are int32 and pair counts uint32.
    static_assert(PT_BATCH_MAX_SLOTS == 0x7FFFFFFF / 3 && PT_BATCH_SLOT_TILE == kSweepTile, "pt_batch.h limits");
    if (P > PT_BATCH_MAX_SLOTS || kwd > 0x7FFFFFFF || kwe > 0x7FFFFFFF)
        return fail(PT_EINVAL, "batch of %lld agent slots exceeds the supported total (%d)", (long long)P,
                    PT_BATCH_MAX_SLOTS);
    b->P = P;
    b->n_agents = n_agents;
    b->ntiles = tiles;
    b->K_wd = (uint32_t)kwd;
    b->K_we = (uint32_t)kwe;
    std::vector<int32_t> trep(tiles);
    for (int r = 0; r < R; ++r)
        for (int t = 0; t < b->reps[r].ntiles; ++t) trep[b->reps[r].tile0 + t] = r;
    const int64_t kmax = std::max<int64_t>(kwd, kwe);
    const int64_t n3 = 3 * P + 4 * (int64_t)R;  // segment starts are rounded up to multiples of 4
    const int64_t nsel = P / kSelAgentsWe + 1;
    if (dev_alloc(&b->tile_rep, tiles) || dev_alloc(&b->st_p, P) || dev_alloc(&b->st_s, P) ||
        dev_alloc(&b->dy_p, P) || dev_alloc(&b->dy_s, P) || dev_alloc(&b->gn_p, P) || dev_alloc(&b->gn_s, P) ||
        dev_alloc(&b->age, P) || dev_alloc(&b->snap, 6 * P) || dev_alloc(&b->household, P) ||
        dev_alloc(&b->hh_off, HH) || dev_alloc(&b->wp_off, WW) || dev_alloc(&b->wp_people, P) ||
        dev_alloc(&b->claim_p, P) || dev_alloc(&b->claim_s, P) || dev_alloc(&b->list_p, P) ||
        dev_alloc(&b->list_s, P) || dev_alloc(&b->infected, P) || dev_alloc(&b->staging, P) ||
        dev_alloc(&b->blk_counts, (int64_t)tiles * 16) || dev_alloc(&b->blk_ninf, tiles) ||
        dev_alloc(&b->blk_off, tiles) || dev_alloc(&b->mark, kmax / 32 + 1) || dev_alloc(&b->stage_keys, n3) ||
        dev_alloc(&b->stage_ids, n3) || dev_alloc(&b->sel_keys, n3) || dev_alloc(&b->sel_ids, n3) ||
        dev_alloc(&b->sblk_cnt, nsel) || dev_alloc(&b->sblk_off, nsel) || dev_alloc(&b->er_people, n3) ||
        dev_alloc(&b->er_off, kmax + 1) || dev_alloc(&b->d_counts, (int64_t)R * b->D * 16) ||
        dev_alloc(&b->d_inf, (int64_t)R * b->D) || dev_alloc(&b->seg_blk, kSegBlocks) ||
        dev_alloc(&b->seg_blk0, R) || dev_alloc(&b->seg_nblk, R) || dev_alloc(&b->seg_kb_wd, R + 1) ||
        dev_alloc(&b->seg_kb_we, R + 1) || dev_alloc(&b->seg_K_wd, R) || dev_alloc(&b->seg_K_we, R) ||
        dev_alloc(&b->seg_start, R) || dev_alloc(&b->seg_count, R)) {
        const std::string msg = pt_last_error();
        free_batch_population(b);
        return fail(PT_ENOMEM, "%s", msg.c_str());
    }
    if (table_scratch_alloc(b->tables, std::max<int64_t>(n3, Hmax + 1)) != 0) {
        free_batch_population(b);
        return fail(PT_ENOMEM, "table scratch allocation failed");
    }
    b->tables.wide_bits = b->launch[PT_LAUNCH_SORT_WIDE_BITS];
    b->tables.max_bits = b->launch[PT_LAUNCH_SORT_MAX_BITS];
    HIP_TRY(hipMemcpyAsync(b->tile_rep, trep.data(), sizeof(int32_t) * tiles, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b->d_models, b->models.data(), sizeof(Model) * R, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b->d_reps, b->reps.data(), sizeof(RepInfo) * R, hipMemcpyHostToDevice, st));
    // Segmented tables (used unless PT_LAUNCH_SEGMENTED_TABLES is 0): radix
    // blocks per replicate in proportion to its agents (at least one each),
    // and each replicate's bucket range.
    b->seg_tables = R <= kSegBlocks;
    if (b->seg_tables) {
        std::vector<int32_t> nbr(R), blk0(R), blk(kSegBlocks), Kwd(R), Kwe(R);
        std::vector<uint32_t> kbwd(R + 1), kbwe(R + 1);
        int used = 0;
        for (int r = 0; r < R; ++r) {
            nbr[r] = std::max<int32_t>(1, (int32_t)((double)kSegBlocks * b->reps[r].N / (double)n_agents));
            used += nbr[r];
        }
        while (used > kSegBlocks) {  // take from the replicate with the most blocks
            const int r = (int)(std::max_element(nbr.begin(), nbr.end()) - nbr.begin());
            nbr[r]--;
            used--;
        }
        while (used < kSegBlocks) {  // give to the replicate with the most agents per block
            int best = 0;
            for (int r = 1; r < R; ++r)
                if ((double)b->reps[r].N * nbr[best] > (double)b->reps[best].N * nbr[r]) best = r;
            nbr[best]++;
            used++;
        }
        int acc = 0;
        for (int r = 0; r < R; ++r) {
            Kwd[r] = b->models[r].wd_K;
            Kwe[r] = kWeHours * b->models[r].we_E;
            kbwd[r] = b->reps[r].kb_wd;
            kbwe[r] = b->reps[r].kb_we;
        }
        kbwd[R] = b->K_wd;
        kbwe[R] = b->K_we;
        b->K_local_wd = *std::max_element(Kwd.begin(), Kwd.end());
        b->K_local_we = *std::max_element(Kwe.begin(), Kwe.end());
        for (int r = 0; r < R; ++r) {
            blk0[r] = acc;
            for (int j = 0; j < nbr[r]; ++j) blk[acc + j] = r;
            acc += nbr[r];
        }
        HIP_TRY(hipMemcpyAsync(b->seg_blk, blk.data(), sizeof(int32_t) * kSegBlocks, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(b->seg_blk0, blk0.data(), sizeof(int32_t) * R, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(b->seg_nblk, nbr.data(), sizeof(int32_t) * R, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(b->seg_kb_wd, kbwd.data(), sizeof(uint32_t) * (R + 1), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(b->seg_kb_we, kbwe.data(), sizeof(uint32_t) * (R + 1), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(b->seg_K_wd, Kwd.data(), sizeof(int32_t) * R, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(b->seg_K_we, Kwe.data(), sizeof(int32_t) * R, hipMemcpyHostToDevice, st));
        HIP_TRY(hipStreamSynchronize(st));  // the host vectors above are about to go
    }
    // Padding slots keep status 0 and are never read; zero them for tidiness.
    HIP_TRY(hipMemsetAsync(b->st_p, 0, (size_t)P, st));
    HIP_TRY(hipMemsetAsync(b->st_s, 0, (size_t)P, st));
    // K3-K7 of every replicate into its slices (sel_keys is the scratch).
    b->tables.grouped_keys = false;
    for (int r = 0; r < R; ++r) {
        const RepInfo& ri = b->reps[r];
        const int64_t a = ri.abase;
        const PopArrays Pa{b->hh_types + ri.hbase, b->hh_off + ri.hbase, b->household + a, b->age + a,
                           b->wp_off + ri.wbase, b->wp_people + a, b->st_p + a, b->st_s + a, b->dy_p + a, b->dy_s + a,
                           b->gn_p + a, b->gn_s + a, b->claim_p + a, b->claim_s + a};
        hipError_t e = build_population(b->d_models + r, ri.N, ri.H, ri.W, Pa, b->sel_keys, b->tables,
                                        ri.workplace_off, st);
        if (e != hipSuccess) return fail(PT_EDEVICE, "replicate %d setup: %s", r, hipGetErrorString(e));
    }
    HIP_TRY(hipStreamSynchronize(st));
    // Day tables hold the replicates' key ranges side by side.
    b->tables.grouped_keys = true;
    b->setup_done = true;
    if (reset_batch_state(b)) return PT_EDEVICE;
    if (n_people_out)
        for (int r = 0; r < R; ++r) n_people_out[r] = b->reps[r].N;
    return PT_OK;
}

int pt_batch_seed_infections(pt_batch* b, const int32_t* ids, const int32_t* n_pandemic, const int32_t* n_seasonal) {
    if (!b || !n_pandemic || !n_seasonal) return fail(PT_EINVAL, "null argument");
    if (!b->setup_done) return fail(PT_ESTATE, "seed_infections before setup");
    int64_t total = 0;
    for (int r = 0; r < b->R; ++r) {
        if (n_pandemic[r] < 0 || n_seasonal[r] < 0) return fail(PT_EINVAL, "replicate %d: bad initial infected", r);
        const int n = n_pandemic[r] + n_seasonal[r];
        std::vector<int32_t> sorted(ids + total, ids + total + n);
        for (int32_t v : sorted)
            if (v < 0 || v >= b->reps[r].N) return fail(PT_EINVAL, "replicate %d: initial infected id %d out of range", r, v);
        std::sort(sorted.begin(), sorted.end());
        for (int i = 1; i < n; ++i)
            if (sorted[i] == sorted[i - 1]) return fail(PT_EINVAL, "replicate %d: duplicate initial infected id %d", r, sorted[i]);
        total += n;
    }
    if (total > 0 && !ids) return fail(PT_EINVAL, "null ids");
    HIP_TRY(hipSetDevice(b->device));
    if (total > 0) {
        int32_t* d_ids = nullptr;
        if (dev_alloc(&d_ids, total)) return PT_ENOMEM;
        HIP_TRY(hipMemcpyAsync(d_ids, ids, (size_t)total * sizeof(int32_t), hipMemcpyHostToDevice, b->stream));
        int64_t o = 0;
        for (int r = 0; r < b->R; ++r) {
            const int n = n_pandemic[r] + n_seasonal[r];
            const int64_t a = b->reps[r].abase;
            if (n > 0)
                k_seed_infections<<<(n + 255) / 256, 256, 0, b->stream>>>(d_ids + o, n_pandemic[r], n_seasonal[r],
                                                                          b->st_p + a, b->st_s + a, b->dy_p + a,
                                                                          b->dy_s + a, b->gn_p + a, b->gn_s + a);
            o += n;
        }
        hipError_t e = hipStreamSynchronize(b->stream);
        (void)hipFree(d_ids);
        if (e != hipSuccess) return fail(PT_EDEVICE, "seed infections: %s", hipGetErrorString(e));
    }
    return reset_batch_state(b);
}

int pt_batch_get_info(pt_batch* b, int64_t out[4]) {
    if (!b || !out) return fail(PT_EINVAL, "null argument");
    out[0] = b->R;
    out[1] = b->D;
    out[2] = b->n_agents;
    out[3] = b->P;
    return PT_OK;
}

int pt_batch_run(pt_batch* b, int32_t* counts_out, int32_t* infected_out, int32_t* days_out) {
    if (!b) return fail(PT_EINVAL, "null handle");
    if (!b->setup_done) return fail(PT_ESTATE, "run before setup");
    HIP_TRY(hipSetDevice(b->device));
    DayState g;
    HIP_TRY(hipMemcpyAsync(&g, b->d_g, sizeof g, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (g.day != 0) return fail(PT_ESTATE, "run after run: restore the snapshot first");
    if (b->use_graphs && (!b->day_graph[0] || b->graph_table_mode != b->table_mode))
        if (capture_day_graphs(b)) return PT_EDEVICE;
    HIP_TRY(hipEventRecord(b->run_begin, b->stream));
    // Days go out in groups; after each group the batch's done flag (every
    // replicate has ended) is copied back, and once the group before the one
    // just enqueued had ended the batch no further group is enqueued.  Days
    // after the end are no-ops (every kernel checks the flag).
    constexpr int kRunGroup = 4;
    for (int d = 0, grp = 0; d < b->D; ++grp) {
        const int end = std::min(d + kRunGroup, b->D);
        for (; d < end; ++d) {
            if (b->use_graphs)
                HIP_TRY(hipGraphLaunch(b->day_graph[(d % 7) >= 5 ? 1 : 0], b->stream));
            else if (enqueue_batch_day(b, d))
                return PT_EDEVICE;
        }
        HIP_TRY(hipMemcpyAsync(&b->h_done[grp & 1], &b->d_g->done, sizeof(int32_t), hipMemcpyDeviceToHost,
                               b->stream));
        HIP_TRY(hipEventRecord(b->group_ev[grp & 1], b->stream));
        if (grp > 0) {
            HIP_TRY(hipEventSynchronize(b->group_ev[(grp - 1) & 1]));
            if (b->h_done[(grp - 1) & 1]) break;
        }
    }
    HIP_TRY(hipEventRecord(b->run_end, b->stream));
    HIP_TRY(hipEventSynchronize(b->run_end));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, b->run_begin, b->run_end));
    b->last_run_ms = ms;
    std::vector<DayState> h(b->R);
    HIP_TRY(hipMemcpy(h.data(), b->d_states, sizeof(DayState) * b->R, hipMemcpyDeviceToHost));
    for (int r = 0; r < b->R; ++r)
        if (h[r].fault) return fail(PT_EINVAL, "replicate %d: a kernel saw an errand destination outside the model's range", r);
    if (counts_out) HIP_TRY(hipMemcpy(counts_out, b->d_counts, sizeof(int32_t) * 16 * b->R * b->D, hipMemcpyDeviceToHost));
    if (infected_out) HIP_TRY(hipMemcpy(infected_out, b->d_inf, sizeof(int32_t) * b->R * b->D, hipMemcpyDeviceToHost));
    if (days_out)
        for (int r = 0; r < b->R; ++r) days_out[r] = h[r].days_run;
    return PT_OK;
}

int pt_batch_generation_offsets(pt_batch* b, int32_t r, int32_t* c_p, int32_t* c_s, int32_t len) {
    if (!b || !c_p || !c_s) return fail(PT_EINVAL, "null argument");
    if (!b->setup_done) return fail(PT_ESTATE, "generation offsets before setup");
    if (r < 0 || r >= b->R) return fail(PT_EINVAL, "replicate %d out of range", r);
    const int md = b->params[r].max_days;
    if (len < md + 1) return fail(PT_EINVAL, "len must be >= max_days + 1");
    HIP_TRY(hipSetDevice(b->device));
    const RepInfo& ri = b->reps[r];
    HIP_TRY(hipMemsetAsync(b->d_scratch64, 0, 512 * sizeof(unsigned long long), b->stream));
    k_gen_hist<<<grid_for(ri.N, 256, 1024), 256, 0, b->stream>>>(b->gn_p + ri.abase, b->gn_s + ri.abase, ri.N,
                                                                   b->d_scratch64);
    unsigned long long hist[512];
    HIP_TRY(hipMemcpyAsync(hist, b->d_scratch64, sizeof hist, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    int64_t ap = 0, as = 0;
    for (int k = 0; k < md; ++k) {
        c_p[k] = (int32_t)ap;
        c_s[k] = (int32_t)as;
        ap += (int64_t)hist[k];
        as += (int64_t)hist[256 + k];
    }
    c_p[md] = ri.N;
    c_s[md] = ri.N;
    return PT_OK;
}

int64_t pt_batch_array_bytes(pt_batch* b, int32_t r, int32_t which) {
    if (!b || !b->setup_done || r < 0 || r >= b->R) return -1;
    const RepInfo& ri = b->reps[r];
    switch (which) {
        case PT_STATUS_P: case PT_STATUS_S: case PT_DAY_P: case PT_DAY_S: case PT_GEN_P: case PT_GEN_S:
        case PT_AGE: return ri.N;
        case PT_HOUSEHOLD: case PT_WP_PEOPLE: return 4 * (int64_t)ri.N;
        case PT_HH_OFFSETS: return 4 * ((int64_t)ri.H + 1);
        case PT_WP_OFFSETS: return 4 * ((int64_t)ri.W + 1);
        case PT_HH_TYPES: return (int64_t)ri.H + 1;
    }
    return -1;
}

int pt_batch_copy_array(pt_batch* b, int32_t r, int32_t which, void* dst, int64_t bytes) {
    if (!b || !dst) return fail(PT_EINVAL, "null argument");
    if (!b->setup_done) return fail(PT_ESTATE, "copy before setup");
    const int64_t need = pt_batch_array_bytes(b, r, which);
    if (need < 0) return fail(PT_EINVAL, "array %d of replicate %d is not available", which, r);
    if (bytes < need) return fail(PT_EINVAL, "destination holds %lld bytes, %lld needed", (long long)bytes, (long long)need);
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    const RepInfo& ri = b->reps[r];
    const int64_t a = ri.abase;
    const void* src = nullptr;
    switch (which) {
        case PT_STATUS_P: src = b->st_p + a; break;
        case PT_STATUS_S: src = b->st_s + a; break;
        case PT_DAY_P: src = b->dy_p + a; break;
        case PT_DAY_S: src = b->dy_s + a; break;
        case PT_GEN_P: src = b->gn_p + a; break;
        case PT_GEN_S: src = b->gn_s + a; break;
        case PT_AGE: src = b->age + a; break;
        case PT_HOUSEHOLD: src = b->household + a; break;
        case PT_WP_PEOPLE: src = b->wp_people + a; break;
        case PT_HH_OFFSETS: src = b->hh_off + ri.hbase; break;
        case PT_WP_OFFSETS: src = b->wp_off + ri.wbase; break;
        case PT_HH_TYPES: src = b->hh_types + ri.hbase; break;
    }
    HIP_TRY(hipMemcpy(dst, src, (size_t)need, hipMemcpyDeviceToHost));
    return PT_OK;
}

int pt_batch_set_interventions(pt_batch* b, const pt_interventions* iv) {
    if (!b || !iv) return fail(PT_EINVAL, "null argument");
    if (b->setup_done) return fail(PT_ESTATE, "interventions must be set before setup");
    std::vector<Model> m = b->models;
    for (int r = 0; r < b->R; ++r) {
        const int rc = apply_interventions(&iv[r], m[r]);
        if (rc != PT_OK) return fail(rc, "replicate %d: %s", r, pt_last_error());
    }
    b->models = m;
    return PT_OK;
}

int pt_batch_set_launch(pt_batch* b, int32_t knob, int32_t value) {
    if (!b) return fail(PT_EINVAL, "null handle");
    if (int rc = check_launch(knob, value)) return rc;
    b->launch[knob] = value;
    b->grid_contacts = b->cus * b->launch[PT_LAUNCH_CONTACT_BLOCKS_PER_CU];
    b->tables.wide_bits = b->launch[PT_LAUNCH_SORT_WIDE_BITS];
    b->tables.max_bits = b->launch[PT_LAUNCH_SORT_MAX_BITS];
    drop_day_graphs(b);  // captured with the old launch shapes
    return PT_OK;
}

int pt_batch_get_launch(pt_batch* b, int32_t knob) {
    if (!b) return fail(PT_EINVAL, "null handle");
    if (knob < 0 || knob >= PT_LAUNCH_KNOBS) return fail(PT_EINVAL, "launch setting %d unknown", knob);
    return b->launch[knob];
}

int pt_batch_use_graphs(pt_batch* b, int32_t enable) {
    if (!b) return fail(PT_EINVAL, "null handle");
    b->use_graphs = enable != 0;
    return PT_OK;
}

int pt_batch_set_table_mode(pt_batch* b, int32_t mode) {
    if (!b) return fail(PT_EINVAL, "null handle");
    if (mode < TABLES_AUTO || mode > TABLES_SUBSET) return fail(PT_EINVAL, "table mode %d unknown", mode);
    b->table_mode = mode;
    return PT_OK;
}

int pt_batch_snapshot(pt_batch* b) {
    if (!b) return fail(PT_EINVAL, "null handle");
    if (!b->setup_done) return fail(PT_ESTATE, "snapshot before setup");
    HIP_TRY(hipSetDevice(b->device));
    uint8_t* arr[6] = {b->st_p, b->st_s, b->dy_p, b->dy_s, b->gn_p, b->gn_s};
    for (int i = 0; i < 6; ++i)
        k_copy16<<<grid_for(b->P / 16, 256, 8192), 256, 0, b->stream>>>(
            reinterpret_cast<const uint4*>(arr[i]), reinterpret_cast<uint4*>(b->snap + i * b->P), b->P / 16);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->have_snapshot = true;
    return PT_OK;
}

int pt_batch_restore(pt_batch* b) {
    if (!b) return fail(PT_EINVAL, "null handle");
    if (!b->have_snapshot) return fail(PT_ESTATE, "restore without snapshot");
    HIP_TRY(hipSetDevice(b->device));
    uint8_t* arr[6] = {b->st_p, b->st_s, b->dy_p, b->dy_s, b->gn_p, b->gn_s};
    for (int i = 0; i < 6; ++i)
        k_copy16<<<grid_for(b->P / 16, 256, 8192), 256, 0, b->stream>>>(
            reinterpret_cast<const uint4*>(b->snap + i * b->P), reinterpret_cast<uint4*>(arr[i]), b->P / 16);
    kb_fill_u64<<<grid_for(b->P, 256, 65535), 256, 0, b->stream>>>(b->claim_p, b->P, ~0ull);
    kb_fill_u64<<<grid_for(b->P, 256, 65535), 256, 0, b->stream>>>(b->claim_s, b->P, ~0ull);
    HIP_TRY(hipGetLastError());
    return reset_batch_state(b);
}

int pt_batch_synchronize(pt_batch* b) {
    if (!b) return fail(PT_EINVAL, "null handle");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return PT_OK;
}

double pt_batch_last_run_ms(pt_batch* b) { return b ? b->last_run_ms : -1.0; }

void pt_batch_destroy(pt_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_day_graphs(b);
    free_batch_population(b);
    dev_free(b->d_models);
    dev_free(b->d_reps);
    dev_free(b->d_states);
    dev_free(b->d_g);
    dev_free(b->rep_I);
    dev_free(b->inf_start);
    dev_free(b->d_scratch64);
    if (b->run_begin) (void)hipEventDestroy(b->run_begin);
    if (b->run_end) (void)hipEventDestroy(b->run_end);
    for (auto e : b->group_ev)
        if (e) (void)hipEventDestroy(e);
    if (b->h_done) (void)hipHostFree(b->h_done);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // extern "C"
