// pt_sim.hip — gfx950 kernels of the PandemicThrust daily loop and the C ABI
// of include/pandemic/pt_abi.h.
//
// Citations are relative to PandemicThrust/PandemicThrust/ of the reference.
//
// Device data (one copy per handle, structure of arrays in HBM):
//   per agent   status/day/generation per strain and age (uint8),
//               household id (int32); claim words per strain (uint64)
//   tables      household offsets, workplace offsets + people, the day's
//               errand offsets (dense bucket layout) + people
//   day state   the reference's host counters (current day, rand_offset,
//               saved errand offset, infected count) kept on the device so
//               that a whole run is enqueued without host round trips.
//
// Infection resolution.  The reference applies infections with a racy
// atomicCAS (PandemicSim.cu:3048-3067).  Here every successful attempt on a
// victim that is susceptible at the start of the day posts
// ((pos*8 + slot) << 32 | gen << 8 | profile) with a 64-bit atomicMin into
// the victim's claim word for that strain; the smallest (infected-list
// position, contact slot) wins, which is the outcome of processing infectors
// in ascending position and contacts in slot order.  The first poster of a
// claim word appends the victim to the strain's claim list; k_apply then
// writes status/day/generation from the winning payload and clears the word.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pandemic/pt_abi.h"
#include "pt_day.h"
#include "pt_model.h"
#include "pt_sort.h"

using namespace pt;

namespace pt {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int check_launch(int32_t knob, int32_t value) {
    if (knob < 0 || knob >= PT_LAUNCH_KNOBS) return fail(PT_EINVAL, "launch setting %d unknown", knob);
    const int lo[PT_LAUNCH_KNOBS] = {1, 1, 1, 0, 0, 4, 0}, hi[PT_LAUNCH_KNOBS] = {64, 64, 9, 1000, 100, 11, 1};
    if (value < lo[knob] || value > hi[knob])
        return fail(PT_EINVAL, "launch setting %d: value %d outside [%d, %d]", knob, value, lo[knob], hi[knob]);
    return PT_OK;
}

}  // namespace pt

namespace {


// ============================ setup kernels ============================

// K1/K2: household types, four per Threefry counter g = 0..H/4
// (hh_peopleCount_functor PandemicSim.cu:3579-3605; kernel_householdTypeAssignment :2329-2357).
__global__ __launch_bounds__(256) void k_household_types(const Model* __restrict__ Mp, uint8_t* __restrict__ types,
                                                         unsigned long long* __restrict__ people) {
    const Model& M = *Mp;
    const int H = M.H;
    __shared__ uint32_t s_w[4];
    uint32_t cnt = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g <= H / 4; g += stride) {
        const U4 u = draw4((uint64_t)g, M.key0, M.key1);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t h = 4 * g + i;
            if (h < H) {
                const int t = fish_household(M, u.w[i]);
                types[h] = (uint8_t)t;
                cnt += (uint32_t)(M.hh_adults[t] + M.hh_children[t]);
            }
        }
    }
    cnt = wave_sum(cnt);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) s_w[wave] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < (int)(blockDim.x / 64); ++w) t += s_w[w];
        atomicAdd(people, t);
    }
}

__global__ void k_household_sizes(const Model* __restrict__ Mp, const uint8_t* __restrict__ types,
                                  uint32_t* __restrict__ sizes) {
    const Model& M = *Mp;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; h <= M.H; h += stride) {
        if (h < M.H) {
            const int t = types[h];
            sizes[h] = (uint32_t)(M.hh_adults[t] + M.hh_children[t]);
        } else {
            sizes[h] = 0;
        }
    }
}

// K4: kernel_generateHouseholds (PandemicSim.cu:2490-2531): adults first.
__global__ void k_generate_households(const Model* __restrict__ Mp, const uint8_t* __restrict__ types,
                                      const int32_t* __restrict__ hh_off, int32_t* __restrict__ household,
                                      uint8_t* __restrict__ age) {
    const Model& M = *Mp;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; h < M.H; h += stride) {
        const int t = types[h];
        const int off = hh_off[h];
        const int na = M.hh_adults[t], nc = M.hh_children[t];
        for (int a = 0; a < na; ++a) {
            household[off + a] = (int32_t)h;
            age[off + a] = kAgeAdult;
        }
        for (int c = 0; c < nc; ++c) {
            household[off + na + c] = (int32_t)h;
            age[off + na + c] = 0;
        }
    }
}

// K5: assignWorkplaceFunctor (PandemicSim.cu:3750-3778), four people per
// counter workplace_off + g, g = 0..N/4.
__global__ void k_assign_workplaces(const Model* __restrict__ Mp, uint64_t workplace_off, uint8_t* __restrict__ age,
                                    uint32_t* __restrict__ wp_keys) {
    const Model& M = *Mp;
    const int N = M.N;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g <= N / 4; g += stride) {
        const U4 u = draw4(workplace_off + (uint64_t)g, M.key0, M.key1);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t p = 4 * g + i;
            if (p < N) {
                uint8_t a = age[p];
                const int32_t wp = assign_workplace(M, u.w[i], &a);
                age[p] = a;
                wp_keys[p] = (uint32_t)wp;
            }
        }
    }
}

// K7: setup_initializeStatusArrays (PandemicSim.cu:3654-3664).  Agents
// vaccinated against a strain (intervention stream draw below the coverage;
// coverage 0 in the reference) start recovered for it.
__global__ void k_init_status(const Model* __restrict__ Mp, int64_t N, uint8_t* st_p, uint8_t* st_s, uint8_t* dy_p,
                              uint8_t* dy_s, uint8_t* gn_p, uint8_t* gn_s) {
    const Model& M = *Mp;
    const bool vacc = M.vacc[0] > 0.f || M.vacc[1] > 0.f;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        uint8_t sp = kSusceptible, ss = kSusceptible;
        if (vacc) {
            const U4 x = draw4(vaccination_counter(i), M.ikey0, M.ikey1);
            if (unit_float(x.w[0]) < M.vacc[0]) sp = kRecovered;
            if (unit_float(x.w[1]) < M.vacc[1]) ss = kRecovered;
        }
        st_p[i] = sp;
        st_s[i] = ss;
        dy_p[i] = kDayNone;
        dy_s[i] = kDayNone;
        gn_p[i] = kGenNone;
        gn_s[i] = kGenNone;
    }
}

__global__ void k_fill_u64(uint64_t* p, int64_t n, uint64_t v) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// ============================ daily kernels ============================

// D1 + D2 fused over the whole population (sweep_block).
__global__ __launch_bounds__(kSweepThreads) void k_sweep(const Model* __restrict__ Mp, int64_t N,
                                                         DayState* __restrict__ st,
                                                         uint8_t* __restrict__ st_p, uint8_t* __restrict__ st_s,
                                                         const uint8_t* __restrict__ dy_p,
                                                         const uint8_t* __restrict__ dy_s,
                                                         uint32_t* __restrict__ blk_counts,
                                                         uint32_t* __restrict__ blk_ninf,
                                                         int32_t* __restrict__ staging) {
    if (st->done) return;
    const int64_t b = blockIdx.x;
    sweep_block(*Mp, N, st->day, b * kSweepTile, st_p, st_s, dy_p, dy_s, blk_counts + b * 16, blk_ninf + b,
                staging + b * kSweepTile, 0);
}

// Day plan (one block): daily counts, the infected count and block offsets
// of the compaction, the stop rule (PandemicSim.cu:899-901) and the RNG
// offsets of the day (:3030-3040 weekday, :1409-1415 weekend, :1278-1279 and
// :1393-1394 contacts).
__global__ __launch_bounds__(kPlanThreads) void k_plan(DayState* __restrict__ st, int nblk,
                                                       const uint32_t* __restrict__ blk_counts,
                                                       const uint32_t* __restrict__ blk_ninf,
                                                       uint32_t* __restrict__ blk_off, int32_t* __restrict__ out_counts,
                                                       int32_t* __restrict__ out_inf, int64_t N, int32_t mark_limit_wd,
                                                       int32_t mark_limit_we, int table_mode) {
    if (st->done) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int NW = kPlanThreads / 64;
    __shared__ uint32_t s_red[NW][16];
    __shared__ uint32_t s_w[NW];
    __shared__ uint32_t s_carry;
    // Counts: each block's 16 counters are one 64-byte row (4 x uint4).
    uint32_t c[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) c[i] = 0;
    const uint4* rows = reinterpret_cast<const uint4*>(blk_counts);
    for (int b = tid; b < nblk; b += kPlanThreads) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint4 v = rows[(int64_t)b * 4 + q];
            c[4 * q] += v.x;
            c[4 * q + 1] += v.y;
            c[4 * q + 2] += v.z;
            c[4 * q + 3] += v.w;
        }
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const uint32_t v = wave_sum(c[i]);
        if (lane == 0) s_red[wave][i] = v;
    }
    // Block offsets: thread tid scans its own run of consecutive blocks, so
    // one block-wide scan of the run totals suffices.
    const int per = (nblk + kPlanThreads - 1) / kPlanThreads;
    const int b0 = tid * per, b1 = min(nblk, b0 + per);
    uint32_t run = 0;
    for (int b = b0; b < b1; ++b) run += blk_ninf[b];
    const uint32_t inc = wave_incl_scan(run, lane);
    if (lane == 63) s_w[wave] = inc;
    __syncthreads();
    const int day = st->day;
    if (tid < 16) {
        uint32_t s = 0;
        for (int w = 0; w < NW; ++w) s += s_red[w][tid];
        out_counts[day * 16 + tid] = (int32_t)s;
    }
    uint32_t wbase = 0, tot = 0;
    for (int w = 0; w < NW; ++w) {
        if (w < wave) wbase += s_w[w];
        tot += s_w[w];
    }
    uint32_t off = wbase + inc - run;
    for (int b = b0; b < b1; ++b) {
        const uint32_t v = blk_ninf[b];
        blk_off[b] = off;
        off += v;
    }
    if (tid == 0) s_carry = tot;
    __syncthreads();
    if (tid == 0) {
        const int32_t I = (int32_t)s_carry;
        st->n_infected = I;
        st->days_run = day + 1;
        out_inf[day] = I;
        if (I == 0) {
            st->done = 1;
            st->skip_full = 1;
            st->skip_subset = 1;
            st->skip_mark = 1;
        } else {
            const bool weekend = (day % 7) >= 5;
            // Queried buckets are marked only when few enough infectors make
            // a subset table likely to pay off; k_decide settles the path.
            const bool mark = table_mode == TABLES_SUBSET ||
                              (table_mode == TABLES_AUTO && I <= (weekend ? mark_limit_we : mark_limit_wd));
            st->skip_mark = mark ? 0 : 1;
            st->skip_full = mark ? 1 : 0;
            st->skip_subset = 1;
            st->errand_off = st->rand_offset;
            st->rand_offset += weekend ? (uint64_t)(2 * N) : (uint64_t)(N / 2);
            st->contact_base = st->rand_offset;
            st->rand_offset += (uint64_t)6 * (uint64_t)I;
        }
    }
}

// Also clears the bucket marks when today marks (k_plan decided).
__global__ void k_compact(const DayState* __restrict__ st, const int32_t* __restrict__ staging,
                          const uint32_t* __restrict__ blk_ninf, const uint32_t* __restrict__ blk_off,
                          int32_t* __restrict__ infected, uint32_t* __restrict__ mark, int64_t nwords) {
    if (st->done) return;
    const uint32_t n = blk_ninf[blockIdx.x];
    const uint32_t off = blk_off[blockIdx.x];
    const int32_t* src = staging + (int64_t)blockIdx.x * kSweepTile;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) infected[off + i] = src[i];
    if (st->skip_mark) return;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += stride) mark[i] = 0u;
}

// D3w: kernel_assignWeekdayAfterschoolAndErrands (PandemicSim.cu:2972-3022).
// Person pair (2g, 2g+1) shares counter errand_off + g.  Keys from
// weekday_keys; hour 0 block [0, N), hour 1 block [N, 2N).
__global__ __launch_bounds__(256) void k_gen_weekday(const Model* __restrict__ Mp, DayState* __restrict__ st,
                                                     const uint8_t* __restrict__ age, uint32_t* __restrict__ keys) {
    if (st->skip_full) return;
    const Model& M = *Mp;
    const int N = M.N;
    const uint64_t off = st->errand_off;
    bool bad = false;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; 2 * g < N; g += stride) {
        const U4 u = draw4(off + (uint64_t)g, M.key0, M.key1);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int64_t p = 2 * g + k;
            if (p < N) {
                const bool adult = age[p] == kAgeAdult;
                int32_t d1, d2;
                weekday_dests(M, u.w[2 * k], u.w[2 * k + 1], adult, d1, d2);
                uint32_t k0, k1;
                bad |= !weekday_keys(M, d1, d2, k0, k1);
                keys[p] = k0;
                keys[N + p] = k1;
            }
        }
    }
    if (bad) st->fault = 1;
}

// D3e: kernel_assignWeekendErrands (PandemicSim.cu:2082-2094); person p uses
// counters errand_off + 2p, +1.  Keys: hour * E + (loc - base), 3 per person.
__global__ __launch_bounds__(256) void k_gen_weekend(const Model* __restrict__ Mp, DayState* __restrict__ st,
                                                     uint32_t* __restrict__ keys) {
    if (st->skip_full) return;
    const Model& M = *Mp;
    const int N = M.N;
    const uint64_t off = st->errand_off;
    const int base = M.we_base, E = M.we_E;
    bool bad = false;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < N; p += stride) {
        int32_t hour[3], loc[3];
        weekend_errands(M, off + (uint64_t)(int64_t)(2 * (int32_t)p), hour, loc);
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            int32_t l = loc[e] - base;
            if ((uint32_t)l >= (uint32_t)E) {
                bad = true;
                l = min(max(l, 0), E - 1);
            }
            keys[3 * p + e] = (uint32_t)(hour[e] * E + l);
        }
    }
    if (bad) st->fault = 1;
}

// ---------------- subset errand tables ----------------

// Chooses today's table path from the number of queried buckets.
__global__ __launch_bounds__(kPlanThreads) void k_decide(DayState* __restrict__ st, const uint32_t* __restrict__ mark,
                                                         int64_t nwords, int32_t K, int forced, int share_pct) {
    if (st->done || st->skip_mark) return;
    __shared__ uint32_t s_w[kPlanThreads / 64];
    uint32_t c = 0;
    for (int64_t i = threadIdx.x; i < nwords; i += kPlanThreads) c += (uint32_t)__popc(mark[i]);
    c = wave_sum(c);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m = 0;
        for (int w = 0; w < kPlanThreads / 64; ++w) m += s_w[w];
        st->n_marked = m;
        const bool subset = forced == TABLES_SUBSET || (forced == TABLES_AUTO && subset_pays(m, K, share_pct));
        st->skip_subset = subset ? 0 : 1;
        st->skip_full = subset ? 1 : 0;
    }
}

// Subset path: stage the pairs of the queried buckets (select_block).
__global__ __launch_bounds__(kSelThreads) void k_select_weekday(const Model* __restrict__ Mp, DayState* __restrict__ st,
                                                                const uint8_t* __restrict__ age,
                                                                const uint32_t* __restrict__ mark,
                                                                uint32_t* __restrict__ stage_keys,
                                                                int32_t* __restrict__ stage_ids,
                                                                uint32_t* __restrict__ blk_cnt) {
    if (st->skip_subset) return;
    const Model& M = *Mp;
    const int64_t b = blockIdx.x;
    if (!select_block<kSelPerWd, false>(M, st->errand_off, age, mark, 0u, false, b * kSelAgentsWd, M.N,
                                        stage_keys + b * kSelStageWd, stage_ids + b * kSelStageWd, blk_cnt + b))
        st->fault = 1;
}

__global__ __launch_bounds__(kSelThreads) void k_select_weekend(const Model* __restrict__ Mp, DayState* __restrict__ st,
                                                                const uint32_t* __restrict__ mark,
                                                                uint32_t* __restrict__ stage_keys,
                                                                int32_t* __restrict__ stage_ids,
                                                                uint32_t* __restrict__ blk_cnt) {
    if (st->skip_subset) return;
    const Model& M = *Mp;
    const int64_t b = blockIdx.x;
    if (!select_block<kSelPerWe, true>(M, st->errand_off, nullptr, mark, 0u, false, b * kSelAgentsWe, M.N,
                                       stage_keys + b * kSelStageWe, stage_ids + b * kSelStageWe, blk_cnt + b))
        st->fault = 1;
}

// Marks the errand buckets today's infectors draw contacts from.
__global__ __launch_bounds__(256) void k_mark_weekday(const Model* __restrict__ Mp, const DayState* __restrict__ st,
                                                      const int32_t* __restrict__ infected, AgentArrays A,
                                                      uint32_t* __restrict__ mark) {
    if (st->done || st->skip_mark) return;
    const Model& M = *Mp;
    const int I = st->n_infected;
    const int day = st->day;
    const uint64_t eoff = st->errand_off;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; pos < I; pos += stride)
        mark_weekday(M, A, day, eoff, infected[pos], mark, 0u);
}

__global__ __launch_bounds__(256) void k_mark_weekend(const Model* __restrict__ Mp, const DayState* __restrict__ st,
                                                      const int32_t* __restrict__ infected, AgentArrays A,
                                                      uint32_t* __restrict__ mark) {
    if (st->done || st->skip_mark) return;
    const Model& M = *Mp;
    const int I = st->n_infected;
    const int day = st->day;
    const uint64_t cbase = st->contact_base;
    const uint64_t eoff = st->errand_off;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; pos < I; pos += stride)
        mark_weekend(M, A, day, cbase, eoff, pos, infected[pos], mark, 0u);
}

// D6w/D6e + D7 over the day's infected list (weekday_contacts,
// weekend_contacts); claims go to the per-strain claim lists.
template <bool WEEKEND>
__global__ __launch_bounds__(kContactThreads) void k_contacts(const Model* __restrict__ Mp, DayState* __restrict__ st,
                                                              const int32_t* __restrict__ infected, AgentArrays A,
                                                              int32_t* __restrict__ list_p,
                                                              int32_t* __restrict__ list_s, uint64_t workplace_off) {
    if (st->done) return;
    const Model& M = *Mp;
    const int I = st->n_infected;
    const int day = st->day;
    const uint64_t cbase = st->contact_base;
    const uint64_t eoff = st->errand_off;
    const int lane = threadIdx.x & 63;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t b0 = (int64_t)blockIdx.x * blockDim.x; b0 < I; b0 += stride) {
        const int64_t pos = b0 + threadIdx.x;
        int32_t vic[8];
        int typ[8];
        bool newp[8], news[8];
        clear_contacts(vic, typ, newp, news);
        if (pos < I) {
            if (WEEKEND)
                weekend_contacts(M, A, day, cbase, eoff, infected[pos], pos, vic, typ, newp, news);
            else
                weekday_contacts(M, A, day, cbase, eoff, workplace_off, infected[pos], pos, vic, typ, newp, news);
        }
        append_claims(st->n_claimed, list_p, list_s, 0, vic, newp, news, lane);
    }
}

// Also logs the day's errand table: full (every pair, n_full) or the subset
// of queried buckets (n_sel pairs).
// tab_pairs logs the pairs of the day's table: n_kept (when set) or n_full
// on full-table days, the selected pairs on subset days.
__global__ void k_end_day(DayState* st, int32_t* __restrict__ tab_full, int32_t* __restrict__ tab_pairs,
                          int32_t n_full, const uint32_t* __restrict__ n_kept) {
    if (st->done) return;
    const int d = st->day;
    tab_full[d] = st->skip_full == 0 ? 1 : 0;
    tab_pairs[d] = st->skip_full == 0 ? (n_kept ? (int32_t)*n_kept : n_full)
                                      : (st->skip_subset == 0 ? (int32_t)st->n_sel : 0);
    st->day += 1;
    st->n_claimed[0] = 0;
    st->n_claimed[1] = 0;
    st->skip_full = 1;
    st->skip_subset = 1;
    if (st->day >= st->max_days) st->done = 1;
}

// Validation trace: which successful attempts won their victim.  A claim
// word holds the smallest (position * 8 + slot) of the day's attempts on
// the victim, the canonical winner (the oracle's sequential order).
// Pandemic winner -> 2, seasonal winner -> 1 (ACTION_INFECT_* codes,
// simParameters.h:23-26).  Runs before k_apply clears the claim words.
__global__ void k_trace_actions(const DayState* __restrict__ st, int C, const int32_t* __restrict__ victim,
                                const uint8_t* __restrict__ attempt, const uint64_t* __restrict__ claim_p,
                                const uint64_t* __restrict__ claim_s, uint8_t* __restrict__ action) {
    if (st->done) return;
    const int64_t n = (int64_t)st->n_infected * C;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t key = (uint32_t)(e / C) * 8u + (uint32_t)(e % C);
        const uint8_t a = attempt[e];
        uint8_t act = 0;
        if ((a & 2) && (uint32_t)(claim_p[victim[e]] >> 32) == key) act += 2;
        if ((a & 1) && (uint32_t)(claim_s[victim[e]] >> 32) == key) act += 1;
        action[e] = act;
    }
}

// Per-agent workplace or school (device_recalcWorkplace, PandemicSim.cu:3872-3890).
__global__ void k_workplace_of(const Model* __restrict__ Mp, uint64_t workplace_off, const uint8_t* __restrict__ age,
                               int64_t N, int32_t* __restrict__ out) {
    const Model& M = *Mp;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const U4 u = draw4(workplace_off + (uint64_t)(i / 4), M.key0, M.key1);
        uint8_t a = age[i];
        out[i] = assign_workplace(M, word(u, (int)(i % 4)), &a);
    }
}

// The day's errand destinations in the reference's encoding hour * W +
// location (people_errands): weekday hour-major [hour 0: N][hour 1: N]
// (PandemicSim.cu:2987-3020), weekend 3 per person (:2082-2094).
__global__ void k_errands_of(const Model* __restrict__ Mp, const DayState* __restrict__ st, int weekend,
                             const uint8_t* __restrict__ age, int64_t N, int32_t* __restrict__ out) {
    const Model& M = *Mp;
    const uint64_t eoff = st->errand_off;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        if (weekend) {
            int32_t hour[3], loc[3];
            weekend_errands(M, eoff + (uint64_t)(2 * i), hour, loc);
            for (int j = 0; j < 3; ++j) out[3 * i + j] = hour[j] * M.W + loc[j];
        } else {
            const U4 ue = draw4(eoff + (uint64_t)(i / 2), M.key0, M.key1);
            const int rs = 2 * (int)(i % 2);
            int32_t d1, d2;
            weekday_dests(M, word(ue, rs), word(ue, rs + 1), age[i] == kAgeAdult, d1, d2);
            out[i] = d1;
            out[N + i] = M.W + d2;
        }
    }
}

// ============================ host side ============================

// The reference model tables (PandemicSim.cu:329-481; simParameters.h).
void base_tables(Model& m) {
    const float hh_cdf[9] = {0.279f, 0.319f, 0.628f, 0.671f, 0.8f, 0.812f, 0.939f, 0.944f, 1.0f};
    const int hh_a[9] = {1, 1, 2, 1, 2, 1, 2, 1, 2};
    const int hh_c[9] = {0, 1, 0, 2, 1, 3, 2, 4, 3};
    const float child_cdf[5] = {0.24f, 0.47f, 0.72f, 0.85f, 1.0f};
    const float wp_pdf[14] = {0.06586f, 0.05802f, 0.30227f, 0.0048f, 0.00997f, 0.203f, 0.09736f,
                              0.10598f, 0.00681f, 0.02599f, 0.f,     0.08749f, 0.03181f, 0.00064f};
    const int wp_max[14] = {3, 3, 3, 2, 2, 3, 3, 2, 2, 2, 0, 2, 2, 2};
    const float shed[6][10] = {
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
    memcpy(m.hh_cdf, hh_cdf, sizeof hh_cdf);
    memcpy(m.hh_adults, hh_a, sizeof hh_a);
    memcpy(m.hh_children, hh_c, sizeof hh_c);
    memcpy(m.child_cdf, child_cdf, sizeof child_cdf);
    memcpy(m.wp_pdf, wp_pdf, sizeof wp_pdf);
    memcpy(m.wp_max, wp_max, sizeof wp_max);
    for (int t = 0; t < kTypes; ++t) m.wd_pdf[t] = m.we_pdf[t] = 0.f;
    m.wd_pdf[9] = 0.61919f;
    m.wd_pdf[11] = 0.27812f;
    m.wd_pdf[12] = 0.06601f;
    m.wd_pdf[13] = 0.03668f;
    m.we_pdf[9] = 0.51493f;
    m.we_pdf[11] = 0.25586f;
    m.we_pdf[12] = 0.1162f;
    m.we_pdf[13] = 0.113f;
    const int wd_assign[4][2] = {{2, 0}, {0, 2}, {1, 1}, {wp_max[kTypeAfterschool], 0}};
    const int we_assign[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
    memcpy(m.wd_assign, wd_assign, sizeof wd_assign);
    memcpy(m.we_assign, we_assign, sizeof we_assign);
    memcpy(m.shed, shed, sizeof shed);
}

const int kBaseCounts[kTypes] = {1000, 7000, 2400, 300, 100, 200, 100, 100, 300, 500, 0, 300, 400, 100};

int round_half_up(double d) { return (int)floor(d + 0.5); }  // roundHalfUp_toInt, PandemicSim.cu:1662

const char* kKernelNames[KC_COUNT] = {"sweep",         "plan",       "compact",   "gen_weekday",
                                      "gen_weekend",   "sort_upsweep", "sort_scan", "sort_downsweep",
                                      "offsets",       "contacts_weekday", "contacts_weekend", "apply",
                                      "end_day",       "mark",       "select",    "subset_sort", "setup"};

// Per-class kernel timing.  mode 0: off; -1: every class; k+1: class k only.
// Launches are also accounted to the simulated day they belong to (day is
// set by the code that enqueues a day; -1 outside the daily loop).
struct EventTimer : LaunchTimer {
    struct Pending {
        int cls;
        size_t start;  // index of the start event (stop = start + 1)
        int day;
    };
    int mode = 0;
    int day = -1;
    std::vector<hipEvent_t> pool;
    std::vector<Pending> pending;
    size_t used = 0;
    double ms[KC_COUNT] = {0};
    int64_t launches[KC_COUNT] = {0};
    std::vector<double> day_ms;  // [day * KC_COUNT + class]
    hipEvent_t next() {
        if (used == pool.size()) {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            pool.push_back(e);
        }
        return pool[used++];
    }
    bool want(int cls) const override { return mode == -1 || mode == cls + 1; }
    void events(int cls, hipEvent_t* start, hipEvent_t* stop) override {
        pending.push_back({cls, used, day});
        *start = next();
        *stop = next();
    }
    // Call after the stream is idle.
    void collect() {
        for (auto& p : pending) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, pool[p.start], pool[p.start + 1]) == hipSuccess) {
                ms[p.cls] += t;
                launches[p.cls] += 1;
                if (p.day >= 0) {
                    if (day_ms.size() < (size_t)(p.day + 1) * KC_COUNT) day_ms.resize((size_t)(p.day + 1) * KC_COUNT, 0.0);
                    day_ms[(size_t)p.day * KC_COUNT + p.cls] += t;
                }
            }
        }
        pending.clear();
        used = 0;
    }
    void reset() {
        pending.clear();
        day_ms.clear();
        day = -1;
        used = 0;
        for (int i = 0; i < KC_COUNT; ++i) {
            ms[i] = 0;
            launches[i] = 0;
        }
    }
    ~EventTimer() {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

}  // namespace

namespace pt {

int build_model(const pt_params* p, const int32_t seed_key[2], Model& m) {
    if (p->max_days < 0 || p->max_days > 255)
        return fail(PT_EINVAL, "max_days %d outside [0,255] (day_t is 8 bits)", p->max_days);
    if (!(p->people_scaling > 0.f) || !(p->location_scaling > 0.f))
        return fail(PT_EINVAL, "scaling factors must be positive");
    if (p->init_pandemic < 0 || p->init_seasonal < 0) return fail(PT_EINVAL, "negative initial infected");
    const double hh = (double)p->people_scaling * (double)1000000;
    if (hh > 2.0e8) return fail(PT_EINVAL, "population scale %g too large", (double)p->people_scaling);
    if ((double)p->location_scaling > 1.0e4) return fail(PT_EINVAL, "location scale too large");
    m = Model{};
    base_tables(m);
    // setup_scaleSimulation (PandemicSim.cu:1127-1162)
    m.H = round_half_up(hh);
    int sum = 0;
    for (int t = 0; t < kTypes; ++t) {
        const int orig = round_half_up((double)kBaseCounts[t]);
        const float prod = p->location_scaling * (float)orig;  // float * int
        int nc = round_half_up((double)prod);
        if (nc == 0 && orig > 0) nc = 1;
        m.wp_count[t] = nc;
        sum += nc;
    }
    m.W = sum;
    int acc = 0;
    for (int t = 0; t < kTypes; ++t) {
        m.wp_offset[t] = acc;
        acc += m.wp_count[t];
    }
    // setup_calculateInfectionData (PandemicSim.cu:2828-2840)
    for (int i = 3; i < 6; ++i)
        for (int j = 0; j < 10; ++j) m.shed[i][j] = m.shed[i][j] * p->asymp_factor;
    const float r[2] = {p->r_pandemic, p->r_seasonal};
    for (int i = 0; i < 2; ++i) {
        const float d = (1.0f - p->asymp_factor) * p->percent_symptomatic;
        m.factor[i] = r[i] / d;
    }
    m.pct = p->percent_symptomatic;
    m.key0 = (uint64_t)(int64_t)seed_key[0];
    m.key1 = (uint64_t)(int64_t)seed_key[1];
    m.wd_base = m.wp_offset[kTypeAfterschool];
    m.wd_E = m.W - m.wd_base;
    m.we_base = m.wp_offset[kFirstErrandRow];
    m.we_E = m.W - m.we_base;
    // Lowest weekday destination of an adult: the first errand location, or
    // offset - 1 of an errand type without locations (fish_weekday_dest).
    int32_t b1 = m.we_base;
    for (int t = kFirstErrandRow; t < kTypes; ++t)
        if (m.wp_count[t] == 0) b1 = std::min(b1, m.wp_offset[t] - 1);
    m.wd_base1 = std::max(b1, m.wd_base);
    m.wd_E1 = m.W - m.wd_base1;
    m.wd_K = m.wd_E + m.wd_E1 + 1;
    // No interventions: the reference's 10-day infectious window from the
    // day of infection; the intervention stream gets its own key.
    m.dur[0] = m.dur[1] = kCulmination;
    m.ikey0 = m.key0 ^ 0x9E3779B97F4A7C15ull;
    m.ikey1 = m.key1 ^ 0xBF58476D1CE4E5B9ull;
    return PT_OK;
}

int apply_interventions(const pt_interventions* iv, Model& m) {
    for (int k = 0; k < 2; ++k) {
        if (iv->latent_days[k] < 0 || iv->latent_days[k] > 30)
            return fail(PT_EINVAL, "latent_days[%d] = %d outside [0, 30]", k, iv->latent_days[k]);
        if (iv->infectious_days[k] < 1 || iv->infectious_days[k] > kCulmination)
            return fail(PT_EINVAL, "infectious_days[%d] = %d outside [1, %d]", k, iv->infectious_days[k], kCulmination);
        if (!(iv->vaccination[k] >= 0.f && iv->vaccination[k] <= 1.f))
            return fail(PT_EINVAL, "vaccination[%d] outside [0, 1]", k);
    }
    if (iv->school_close_start < 0 || iv->school_close_days < 0 || iv->school_close_start > 255 ||
        iv->school_close_days > 255)
        return fail(PT_EINVAL, "school closure window outside [0, 255]");
    if (!(iv->workplace_reduction >= 0.f && iv->workplace_reduction <= 1.f) ||
        !(iv->errand_reduction >= 0.f && iv->errand_reduction <= 1.f))
        return fail(PT_EINVAL, "contact reductions must lie in [0, 1]");
    for (int k = 0; k < 2; ++k) {
        m.lat[k] = iv->latent_days[k];
        m.dur[k] = iv->infectious_days[k];
        m.vacc[k] = iv->vaccination[k];
    }
    m.school_start = iv->school_close_start;
    m.school_days = iv->school_close_days;
    m.wp_red = iv->workplace_reduction;
    m.er_red = iv->errand_reduction;
    m.drops = (m.school_days > 0 || m.wp_red > 0.f || m.er_red > 0.f) ? 1 : 0;
    return PT_OK;
}

void launch_household_types(const Model* d_model, int H, uint8_t* hh_types, unsigned long long* people,
                            hipStream_t stream) {
    (void)hipMemsetAsync(hh_types, 0, (size_t)H + 1, stream);
    k_household_types<<<grid_for(H / 4 + 1, 256, 65535), 256, 0, stream>>>(d_model, hh_types, people);
}

hipError_t build_population(const Model* d_model, int64_t N, int H, int W, const PopArrays& P, uint32_t* tmp,
                            TableScratch& ts, uint64_t workplace_off, hipStream_t st, const hipEvent_t* marks) {
    if (marks) (void)hipEventRecord(marks[0], st);
    // K3/K4: household offsets by exclusive scan, then people (PandemicSim.cu:2573-2613).
    k_household_sizes<<<grid_for((int64_t)H + 1, 256, 65535), 256, 0, st>>>(d_model, P.hh_types, tmp);
    hipError_t e = exclusive_scan_u32(tmp, (uint32_t*)P.hh_off, (int64_t)H + 1, ts.scan_part, nullptr, st);
    if (e != hipSuccess) return e;
    k_generate_households<<<grid_for(H, 256, 65535), 256, 0, st>>>(d_model, P.hh_types, P.hh_off, P.household, P.age);
    if (marks) (void)hipEventRecord(marks[1], st);
    // K5/K6: workplaces, then the workplace table (PandemicSim.cu:3812-3862).
    k_assign_workplaces<<<grid_for(N / 4 + 1, 256, 65535), 256, 0, st>>>(d_model, workplace_off, P.age, tmp);
    e = build_table(ts, tmp, nullptr, N, nullptr, W, ID_IDENTITY, (int32_t)N, P.wp_people, P.wp_off, nullptr, st,
                    nullptr);
    if (e != hipSuccess) return e;
    if (marks) (void)hipEventRecord(marks[2], st);
    // K7: statuses; claim words start empty.
    k_init_status<<<grid_for(N, 256, 65535), 256, 0, st>>>(d_model, N, P.st_p, P.st_s, P.dy_p, P.dy_s, P.gn_p,
                                                           P.gn_s);
    k_fill_u64<<<grid_for(N, 256, 65535), 256, 0, st>>>(P.claim_p, N, ~0ull);
    k_fill_u64<<<grid_for(N, 256, 65535), 256, 0, st>>>(P.claim_s, N, ~0ull);
    if (marks) (void)hipEventRecord(marks[3], st);
    return hipGetLastError();
}

}  // namespace pt

struct pt_sim {
    int device = 0;
    hipStream_t stream = nullptr;
    pt_params params{};
    Model model{};
    Model* d_model = nullptr;
    int64_t N = 0;
    int H = 0, W = 0;
    bool setup_done = false, seeded = false, have_snapshot = false;
    uint64_t workplace_off = 0, rand_after_setup = 0;
    // agent arrays
    uint8_t *st_p = nullptr, *st_s = nullptr, *dy_p = nullptr, *dy_s = nullptr, *gn_p = nullptr, *gn_s = nullptr;
    uint8_t* age = nullptr;
    uint8_t* snap = nullptr;  // 6N bytes
    int32_t *household = nullptr, *hh_off = nullptr, *wp_off = nullptr, *wp_people = nullptr;
    uint8_t* hh_types = nullptr;
    int32_t *infected = nullptr, *staging = nullptr;
    uint32_t* keys = nullptr;        // 3N
    int32_t* er_people = nullptr;    // 3N
    int32_t* er_off = nullptr;       // dense, max(wd_K, 10 we_E) + 1
    uint64_t *claim_p = nullptr, *claim_s = nullptr;
    int32_t *list_p = nullptr, *list_s = nullptr;
    uint32_t *blk_counts = nullptr, *blk_ninf = nullptr, *blk_off = nullptr;
    uint32_t* mark = nullptr;                                // queried-bucket bitmap
    uint32_t* stage_keys = nullptr;                          // selection staging (per-block runs)
    int32_t* stage_ids = nullptr;
    uint32_t* sel_keys = nullptr;                            // 3N, compacted selection
    int32_t* sel_ids = nullptr;                              // 3N
    uint32_t *sblk_cnt = nullptr, *sblk_off = nullptr;       // per selection block
    int table_mode = TABLES_AUTO;
    int nblk_sweep = 0;
    DayState* d_state = nullptr;
    int32_t *d_counts = nullptr, *d_inf = nullptr;  // [256*16], [256]
    int32_t *d_tab_full = nullptr, *d_tab_pairs = nullptr;  // [256] errand-table log per day
    unsigned long long* d_scratch64 = nullptr;      // 512
    TableScratch tables;
    int host_day = 0;      // days enqueued since setup / restore
    int last_weekend = -1; // layout of the errand table of the last day enqueued
    int64_t last_table_n = 0;
    EventTimer timer;
    hipEvent_t run_begin = nullptr, run_end = nullptr;
    // pt_sim_run's early stop: the loop's `done` flag copied back after each
    // group of days (pinned, two slots) and the groups' completion events.
    int32_t* h_done = nullptr;
    hipEvent_t group_ev[2] = {nullptr, nullptr};
    hipEvent_t setup_ev[6] = {};               // population size (2), build_population marks (4)
    double setup_ms[PT_SETUP_PHASES] = {0};
    double last_run_ms = 0;
    int grid_contacts = 1024, grid_gen = 2048;
    // Validation trace of the last contact day (pt_sim_enable_trace).
    bool trace_on = false;
    int32_t *tr_victim = nullptr, *tr_loc = nullptr;
    uint8_t *tr_type = nullptr, *tr_attempt = nullptr, *tr_action = nullptr, *tr_draw = nullptr, *tr_prof = nullptr;
    int cus = 128;                                   // compute units of the device
    int launch[PT_LAUNCH_KNOBS] = {8, 16, 8, 100, 75, 8, 1};  // pt_launch_knob values
};

namespace {

double ms_between(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

double event_ms(hipEvent_t a, hipEvent_t b) {
    float t = 0.f;
    return hipEventElapsedTime(&t, a, b) == hipSuccess ? (double)t : 0.0;
}

// Buffers sized by the population (allocated in pt_sim_setup).
void free_trace(pt_sim* s) {
    dev_free(s->tr_victim);
    dev_free(s->tr_type);
    dev_free(s->tr_loc);
    dev_free(s->tr_attempt);
    dev_free(s->tr_action);
    dev_free(s->tr_draw);
    dev_free(s->tr_prof);
}

int alloc_trace(pt_sim* s) {
    free_trace(s);
    const int64_t n = 8 * s->N;
    if (dev_alloc(&s->tr_victim, n) || dev_alloc(&s->tr_type, n) || dev_alloc(&s->tr_loc, n) ||
        dev_alloc(&s->tr_attempt, n) || dev_alloc(&s->tr_action, n) || dev_alloc(&s->tr_draw, n) ||
        dev_alloc(&s->tr_prof, n)) {
        free_trace(s);
        return PT_ENOMEM;
    }
    return 0;
}

void free_population(pt_sim* s) {
    free_trace(s);
    dev_free(s->st_p);
    dev_free(s->st_s);
    dev_free(s->dy_p);
    dev_free(s->dy_s);
    dev_free(s->gn_p);
    dev_free(s->gn_s);
    dev_free(s->age);
    dev_free(s->snap);
    dev_free(s->household);
    dev_free(s->hh_off);
    dev_free(s->wp_off);
    dev_free(s->wp_people);
    dev_free(s->hh_types);
    dev_free(s->infected);
    dev_free(s->staging);
    dev_free(s->keys);
    dev_free(s->er_people);
    dev_free(s->er_off);
    dev_free(s->claim_p);
    dev_free(s->claim_s);
    dev_free(s->list_p);
    dev_free(s->list_s);
    dev_free(s->blk_counts);
    dev_free(s->blk_ninf);
    dev_free(s->blk_off);
    dev_free(s->mark);
    dev_free(s->stage_keys);
    dev_free(s->stage_ids);
    dev_free(s->sel_keys);
    dev_free(s->sel_ids);
    dev_free(s->sblk_cnt);
    dev_free(s->sblk_off);
    table_scratch_free(s->tables);
    s->setup_done = s->seeded = s->have_snapshot = false;
}

AgentArrays agent_arrays(pt_sim* s) {
    AgentArrays A;
    A.age = s->age;
    A.household = s->household;
    A.hh_off = s->hh_off;
    A.wp_off = s->wp_off;
    A.wp_people = s->wp_people;
    A.er_off = s->er_off;
    A.er_people = s->er_people;
    A.st_p = s->st_p;
    A.st_s = s->st_s;
    A.dy_p = s->dy_p;
    A.dy_s = s->dy_s;
    A.gn_p = s->gn_p;
    A.gn_s = s->gn_s;
    A.claim_p = s->claim_p;
    A.claim_s = s->claim_s;
    A.trace_victim = s->trace_on ? s->tr_victim : nullptr;
    A.trace_type = s->trace_on ? s->tr_type : nullptr;
    A.trace_loc = s->trace_on ? s->tr_loc : nullptr;
    A.trace_attempt = s->trace_on ? s->tr_attempt : nullptr;
    A.trace_draw = s->trace_on ? s->tr_draw : nullptr;
    A.trace_prof = s->trace_on ? s->tr_prof : nullptr;
    return A;
}

// Enqueues one loop iteration for host day index d (runToCompletion body,
// PandemicSim.cu:882-954).  Every kernel is a no-op once the device state
// says the loop has ended.
int enqueue_day(pt_sim* s, int d) {
    const bool weekend = (d % 7) >= 5;
    hipStream_t stc = s->stream;
    s->timer.day = d;
    LaunchTimer* tm = &s->timer;
    const int N = (int)s->N;
    DayState* st = s->d_state;
    launch(tm, KC_SWEEP, k_sweep, dim3(s->nblk_sweep), dim3(kSweepThreads), stc, (const Model*)s->d_model, s->N, st,
           s->st_p, s->st_s,
           (const uint8_t*)s->dy_p, (const uint8_t*)s->dy_s, s->blk_counts, s->blk_ninf, s->staging);
    // Marking pays off while the queried share of the buckets is likely to
    // stay below the subset threshold: I <= K * mark_limit / 100.
    const int mpct = s->launch[PT_LAUNCH_MARK_LIMIT];
    const int32_t lim_wd = (int32_t)std::min<int64_t>(mark_limit(s->model.wd_K, mpct), INT32_MAX);
    const int32_t lim_we = (int32_t)std::min<int64_t>(mark_limit((int64_t)kWeHours * s->model.we_E, mpct), INT32_MAX);
    launch(tm, KC_PLAN, k_plan, dim3(1), dim3(kPlanThreads), stc, st, s->nblk_sweep, (const uint32_t*)s->blk_counts,
           (const uint32_t*)s->blk_ninf, s->blk_off, s->d_counts, s->d_inf, s->N, lim_wd, lim_we, s->table_mode);
    const int64_t n = weekend ? 3 * s->N : 2 * s->N;
    const int32_t K = weekend ? kWeHours * s->model.we_E : s->model.wd_K;
    const IdMode mode = weekend ? ID_DIV3 : ID_HOUR_MAJOR;
    // Which buckets will be queried today (marks cleared by the compaction),
    // and which table path to take.
    const int64_t nwords = ((int64_t)K + 31) / 32;
    launch(tm, KC_COMPACT, k_compact, dim3(s->nblk_sweep), dim3(256), stc, (const DayState*)st,
           (const int32_t*)s->staging, (const uint32_t*)s->blk_ninf, (const uint32_t*)s->blk_off, s->infected,
           s->mark, nwords);
    AgentArrays A = agent_arrays(s);
    const Model* M = s->d_model;
    const int32_t* infected = s->infected;
    if (weekend)
        launch(tm, KC_MARK, k_mark_weekend, dim3(s->grid_contacts), dim3(256), stc, M, (const DayState*)st, infected, A,
               s->mark);
    else
        launch(tm, KC_MARK, k_mark_weekday, dim3(s->grid_contacts), dim3(256), stc, M, (const DayState*)st, infected, A,
               s->mark);
    launch(tm, KC_MARK, k_decide, dim3(1), dim3(kPlanThreads), stc, st, (const uint32_t*)s->mark, nwords, K,
           s->table_mode, s->launch[PT_LAUNCH_SUBSET_SHARE]);
    // Full path: every agent's keys, implicit ids.
    if (weekend)
        launch(tm, KC_GEN_WEEKEND, k_gen_weekend, dim3(s->grid_gen), dim3(256), stc, M, st, s->keys);
    else
        launch(tm, KC_GEN_WEEKDAY, k_gen_weekday, dim3(s->grid_gen), dim3(256), stc, M, st, (const uint8_t*)s->age,
               s->keys);
    // Subset path: only pairs of queried buckets, explicit ids.
    const int nb = (int)((s->N + (weekend ? kSelAgentsWe : kSelAgentsWd) - 1) / (weekend ? kSelAgentsWe : kSelAgentsWd));
    if (weekend)
        launch(tm, KC_SELECT, k_select_weekend, dim3(nb), dim3(kSelThreads), stc, M, st, (const uint32_t*)s->mark,
               s->stage_keys, s->stage_ids, s->sblk_cnt);
    else
        launch(tm, KC_SELECT, k_select_weekday, dim3(nb), dim3(kSelThreads), stc, M, st, (const uint8_t*)s->age,
               (const uint32_t*)s->mark, s->stage_keys, s->stage_ids, s->sblk_cnt);
    hipError_t e =
        exclusive_scan_u32(s->sblk_cnt, s->sblk_off, nb, s->tables.scan_part, &st->skip_subset, stc, tm, KC_SELECT);
    if (e != hipSuccess) return fail(PT_EDEVICE, "selection scan: %s", hipGetErrorString(e));
    launch(tm, KC_SELECT, k_compact_runs, dim3(nb), dim3(256), stc, (const int32_t*)&st->skip_subset,
           (const uint32_t*)s->stage_keys, (const int32_t*)s->stage_ids, (const uint32_t*)s->sblk_cnt,
           (const uint32_t*)s->sblk_off, s->sel_keys, s->sel_ids, weekend ? kSelStageWe : kSelStageWd, &st->n_sel, nb);
    // One table build serves both paths: its first pass reads the selected
    // pairs on subset days (skip_full set) and every agent's keys otherwise.
    // Weekday: the last bucket collects the hour-1 keys nobody queries
    // (children stay at their after-school site), left out of the full table
    // (the selection never holds it).
    AltInput sel;
    sel.flag = &st->skip_full;
    sel.keys = s->sel_keys;
    sel.vals = s->sel_ids;
    sel.n_dev = &st->n_sel;
    e = build_table(s->tables, s->keys, nullptr, n, nullptr, K, mode, N, s->er_people, s->er_off, &st->done, stc, tm,
                    weekend ? kNoDrop : (uint32_t)(K - 1), &sel);
    if (e != hipSuccess) return fail(PT_EDEVICE, "table build: %s", hipGetErrorString(e));
    if (weekend)
        launch(tm, KC_CONTACTS_WEEKEND, k_contacts<true>, dim3(s->grid_contacts), dim3(kContactThreads), stc, M, st,
               infected, A, s->list_p, s->list_s, s->workplace_off);
    else
        launch(tm, KC_CONTACTS_WEEKDAY, k_contacts<false>, dim3(s->grid_contacts), dim3(kContactThreads), stc, M, st,
               infected, A, s->list_p, s->list_s, s->workplace_off);
    if (s->trace_on)
        k_trace_actions<<<s->grid_contacts, 256, 0, stc>>>((const DayState*)st, weekend ? 5 : 8,
                                                            (const int32_t*)s->tr_victim,
                                                            (const uint8_t*)s->tr_attempt,
                                                            (const uint64_t*)s->claim_p,
                                                            (const uint64_t*)s->claim_s, s->tr_action);
    launch(tm, KC_APPLY, k_apply, dim3(s->grid_contacts), dim3(256), stc, (const DayState*)st, s->claim_p, s->claim_s,
           (const int32_t*)s->list_p, (const int32_t*)s->list_s, s->st_p, s->st_s, s->dy_p, s->dy_s, s->gn_p, s->gn_s);
    launch(tm, KC_END_DAY, k_end_day, dim3(1), dim3(1), stc, st, s->d_tab_full, s->d_tab_pairs, (int32_t)n,
           weekend ? (const uint32_t*)nullptr : (const uint32_t*)s->tables.n_kept);
    e = hipGetLastError();
    if (e != hipSuccess) return fail(PT_EDEVICE, "day launch: %s", hipGetErrorString(e));
    s->last_weekend = weekend ? 1 : 0;
    s->last_table_n = n;
    s->host_day = d + 1;
    s->timer.day = -1;
    return 0;
}

int reset_day_state(pt_sim* s) {
    DayState h{};
    h.day = 0;
    h.done = s->params.max_days <= 0 ? 1 : 0;
    h.rand_offset = s->rand_after_setup;
    h.max_days = s->params.max_days;
    h.skip_full = 1;
    h.skip_subset = 1;
    HIP_TRY(hipMemcpyAsync(s->d_state, &h, sizeof h, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->host_day = 0;
    s->last_weekend = -1;
    return 0;
}

int read_state(pt_sim* s, DayState* h) {
    HIP_TRY(hipMemcpyAsync(h, s->d_state, sizeof *h, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return 0;
}

// Parity hook: the last weekday table in the reference's layout (bucket
// hour * W + location, lower_bound offsets).  The reference's hour-1 bucket
// of a location in [wd_base, wd_base1) holds exactly the people of its hour-0
// bucket: only children reach those locations, and a child's hour-1
// destination is its hour-0 one (weekday_dests).
int expand_weekday_table(pt_sim* s, int32_t* offsets_out, int32_t* people_out, int64_t people_len) {
    const Model& m = s->model;
    const int W = s->W, E = m.wd_E, E1 = m.wd_E1, A = m.wd_base1 - m.wd_base;
    std::vector<int32_t> dense((size_t)m.wd_K + 1);
    HIP_TRY(hipMemcpy(dense.data(), s->er_off, dense.size() * 4, hipMemcpyDeviceToHost));
    const int32_t h1 = dense[E];            // first hour-1 entry
    const int32_t kids = dense[A] - dense[0];  // hour-0 entries below wd_base1
    const int32_t end1 = dense[E + E1];     // end of the queried hour-1 buckets
    if (offsets_out) {
        for (int l = 0; l < W; ++l) {
            offsets_out[l] = l < m.wd_base ? dense[0] : dense[l - m.wd_base];
            int32_t v;
            if (l < m.wd_base)
                v = h1;
            else if (l < m.wd_base1)
                v = h1 + dense[l - m.wd_base] - dense[0];
            else
                v = dense[E + (l - m.wd_base1)] + kids;
            offsets_out[W + l] = v;
        }
        offsets_out[2 * (int64_t)W] = end1 + kids;
    }
    if (people_out) {
        const int64_t total = (int64_t)end1 + kids;
        if (total > people_len) return fail(PT_EDEVICE, "weekday table holds %lld entries, %lld expected",
                                            (long long)total, (long long)people_len);
        std::vector<int32_t> dev((size_t)end1);
        if (end1 > 0) HIP_TRY(hipMemcpy(dev.data(), s->er_people, (size_t)end1 * 4, hipMemcpyDeviceToHost));
        int32_t* o = people_out;
        o = std::copy(dev.begin(), dev.begin() + h1, o);
        o = std::copy(dev.begin() + dense[0], dev.begin() + dense[A], o);
        o = std::copy(dev.begin() + h1, dev.end(), o);
        std::fill(o, people_out + people_len, 0);
    }
    return PT_OK;
}

}  // namespace

// ============================ C ABI ============================
extern "C" {

const char* pt_last_error(void) { return g_err.c_str(); }

int pt_glibc_seeds(int32_t core_seed, int32_t seeds_out[4]) {
    if (!seeds_out) return fail(PT_EINVAL, "null output");
    srand((unsigned)core_seed);
    for (int i = 0; i < 4; ++i) seeds_out[i] = rand();
    return PT_OK;
}

int pt_glibc_initial_infected(int32_t core_seed, int32_t n_people, int32_t n, int32_t* ids_out) {
    if (n < 0 || n > n_people || (n > 0 && !ids_out)) return fail(PT_EINVAL, "initial infected exceeds population");
    srand((unsigned)core_seed);
    for (int i = 0; i < 4; ++i) (void)rand();
    for (int i = 0; i < n; ++i) {
        bool dup;
        do {
            ids_out[i] = rand() % n_people;
            dup = false;
            for (int j = 0; j < i; ++j)
                if (ids_out[j] == ids_out[i]) {
                    dup = true;
                    break;
                }
        } while (dup);
    }
    return PT_OK;
}

int pt_sim_create(const pt_params* p, const int32_t seed_key[2], int32_t device, pt_sim** out) {
    if (!p || !seed_key || !out) return fail(PT_EINVAL, "null argument");
    *out = nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    Model m{};
    const int rc = build_model(p, seed_key, m);
    if (rc != PT_OK) return rc;
    const auto t1 = std::chrono::steady_clock::now();
    HIP_TRY(hipSetDevice(device));
    pt_sim* s = new pt_sim();
    s->setup_ms[PT_SETUP_SCALE] = ms_between(t0, t1);
    s->device = device;
    s->params = *p;
    s->model = m;
    s->H = m.H;
    s->W = m.W;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete s;
        return fail(PT_EDEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        s->cus = cus;
    s->grid_contacts = s->cus * s->launch[PT_LAUNCH_CONTACT_BLOCKS_PER_CU];
    s->grid_gen = s->cus * s->launch[PT_LAUNCH_GEN_BLOCKS_PER_CU];
    if (dev_alloc(&s->d_model, 1) || dev_alloc(&s->d_state, 1) || dev_alloc(&s->d_counts, 256 * 16) ||
        dev_alloc(&s->d_inf, 256) || dev_alloc(&s->d_scratch64, 512) || dev_alloc(&s->d_tab_full, 256) ||
        dev_alloc(&s->d_tab_pairs, 256)) {
        const std::string msg = g_err;
        pt_sim_destroy(s);
        return fail(PT_ENOMEM, "%s", msg.c_str());
    }
    (void)hipEventCreate(&s->run_begin);
    (void)hipEventCreate(&s->run_end);
    for (auto& e : s->setup_ev) (void)hipEventCreate(&e);
    for (auto& e : s->group_ev) (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
    if (hipHostMalloc((void**)&s->h_done, 2 * sizeof(int32_t), hipHostMallocDefault) != hipSuccess) {
        s->h_done = nullptr;
        pt_sim_destroy(s);
        return fail(PT_ENOMEM, "pinned host allocation failed");
    }
    s->setup_ms[PT_SETUP_PUSH] = ms_between(t1, std::chrono::steady_clock::now());
    *out = s;
    return PT_OK;
}

int pt_sim_setup(pt_sim* s, int64_t* n_people_out) {
    if (!s) return fail(PT_EINVAL, "null handle");
    HIP_TRY(hipSetDevice(s->device));
    free_population(s);
    Model& m = s->model;
    const int H = s->H;
    hipStream_t st = s->stream;
    // K1/K2: household types and the population size (PandemicSim.cu:3634-3651, :2552-2571).
    if (dev_alloc(&s->hh_types, (int64_t)H + 1)) return PT_ENOMEM;
    HIP_TRY(hipMemsetAsync(s->d_scratch64, 0, sizeof(unsigned long long), st));
    m.N = 0;
    HIP_TRY(hipMemcpyAsync(s->d_model, &m, sizeof m, hipMemcpyHostToDevice, st));
    HIP_TRY(hipEventRecord(s->setup_ev[0], st));
    launch_household_types(s->d_model, H, s->hh_types, s->d_scratch64, st);
    HIP_TRY(hipEventRecord(s->setup_ev[1], st));
    unsigned long long npeople = 0;
    HIP_TRY(hipMemcpyAsync(&npeople, s->d_scratch64, sizeof npeople, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    const auto t_sized0 = std::chrono::steady_clock::now();
    if (npeople == 0 || npeople > (unsigned long long)(0x7FFFFFFF / 8))
        return fail(PT_EINVAL, "population %llu outside the supported range", npeople);
    const int64_t N = (int64_t)npeople;
    if ((int64_t)s->params.init_pandemic + s->params.init_seasonal > N)
        return fail(PT_EINVAL, "initial infected exceed the population");
    s->N = N;
    m.N = (int32_t)N;
    HIP_TRY(hipMemcpyAsync(s->d_model, &m, sizeof m, hipMemcpyHostToDevice, st));
    // Per-agent arrays (setup_sizeGlobalArrays, PandemicSim.cu:1047-1107).
    s->nblk_sweep = (int)((N + kSweepTile - 1) / kSweepTile);
    const int64_t kmax = std::max<int64_t>((int64_t)m.wd_K, (int64_t)kWeHours * m.we_E);
    const int64_t n3 = 3 * N;
    const int64_t Npad = ((N + 15) / 16) * 16;  // the sweep reads whole uint4 words
    const int64_t stage_n = std::max<int64_t>((N / kSelAgentsWd + 1) * kSelStageWd, (N / kSelAgentsWe + 1) * kSelStageWe);
    if (dev_alloc(&s->st_p, Npad) || dev_alloc(&s->st_s, Npad) || dev_alloc(&s->dy_p, Npad) ||
        dev_alloc(&s->dy_s, Npad) || dev_alloc(&s->gn_p, Npad) || dev_alloc(&s->gn_s, Npad) ||
        dev_alloc(&s->age, N) || dev_alloc(&s->snap, 6 * Npad) || dev_alloc(&s->household, N) ||
        dev_alloc(&s->hh_off, (int64_t)H + 1) ||
        dev_alloc(&s->wp_off, (int64_t)s->W + 1) || dev_alloc(&s->wp_people, N) || dev_alloc(&s->infected, N) ||
        dev_alloc(&s->staging, (int64_t)s->nblk_sweep * kSweepTile) || dev_alloc(&s->keys, n3) ||
        dev_alloc(&s->er_people, n3) || dev_alloc(&s->er_off, std::max<int64_t>(kmax, (int64_t)s->W) + 1) ||
        dev_alloc(&s->claim_p, N) || dev_alloc(&s->claim_s, N) || dev_alloc(&s->list_p, N) ||
        dev_alloc(&s->list_s, N) || dev_alloc(&s->blk_counts, (int64_t)s->nblk_sweep * 16) ||
        dev_alloc(&s->blk_ninf, s->nblk_sweep) || dev_alloc(&s->blk_off, s->nblk_sweep) ||
        dev_alloc(&s->mark, (kmax + 31) / 32 + 1) || dev_alloc(&s->stage_keys, stage_n) ||
        dev_alloc(&s->stage_ids, stage_n) ||
        dev_alloc(&s->sel_keys, n3) || dev_alloc(&s->sel_ids, n3) || dev_alloc(&s->sblk_cnt, N / kSelAgentsWe + 2) ||
        dev_alloc(&s->sblk_off, N / kSelAgentsWe + 2)) {
        const std::string msg = g_err;
        free_population(s);
        return fail(PT_ENOMEM, "%s", msg.c_str());
    }
    if (table_scratch_alloc(s->tables, n3) != 0) {
        free_population(s);
        return fail(PT_ENOMEM, "table scratch allocation failed");
    }
    s->tables.wide_bits = s->launch[PT_LAUNCH_SORT_WIDE_BITS];
    s->tables.max_bits = s->launch[PT_LAUNCH_SORT_MAX_BITS];
    if (s->trace_on && alloc_trace(s)) {
        free_population(s);
        return fail(PT_ENOMEM, "trace allocation failed");
    }
    const double sizing_ms = ms_between(t_sized0, std::chrono::steady_clock::now());
    // K3-K7 (the keys buffer is the scratch; the rand offset after the
    // households is the workplace counter, PandemicSim.cu:2570-2571, :3817).
    s->workplace_off = (uint64_t)(H / 4);
    s->rand_after_setup = s->workplace_off + (uint64_t)(N / 4);
    const PopArrays P{s->hh_types, s->hh_off, s->household, s->age, s->wp_off, s->wp_people, s->st_p, s->st_s,
                      s->dy_p, s->dy_s, s->gn_p, s->gn_s, s->claim_p, s->claim_s};
    hipError_t e = build_population(s->d_model, N, H, s->W, P, s->keys, s->tables, s->workplace_off, st,
                                    s->setup_ev + 2);
    if (e != hipSuccess) return fail(PT_EDEVICE, "population setup: %s", hipGetErrorString(e));
    HIP_TRY(hipStreamSynchronize(st));
    s->setup_ms[PT_SETUP_POPULATION_SIZE] = event_ms(s->setup_ev[0], s->setup_ev[1]);
    s->setup_ms[PT_SETUP_SIZE_ARRAYS] = sizing_ms;
    s->setup_ms[PT_SETUP_HOUSEHOLDS] = event_ms(s->setup_ev[2], s->setup_ev[3]);
    s->setup_ms[PT_SETUP_WORKPLACES] = event_ms(s->setup_ev[3], s->setup_ev[4]);
    s->setup_ms[PT_SETUP_STATUS] = event_ms(s->setup_ev[4], s->setup_ev[5]);
    s->setup_done = true;
    s->seeded = false;
    if (reset_day_state(s)) return PT_EDEVICE;
    if (n_people_out) *n_people_out = N;
    return PT_OK;
}

int pt_sim_seed_infections(pt_sim* s, const int32_t* ids, int32_t n_p, int32_t n_s) {
    if (!s) return fail(PT_EINVAL, "null handle");
    if (!s->setup_done) return fail(PT_ESTATE, "seed_infections before setup");
    if (n_p < 0 || n_s < 0 || (n_p + n_s > 0 && !ids)) return fail(PT_EINVAL, "bad initial infected");
    const int n = n_p + n_s;
    for (int i = 0; i < n; ++i)
        if (ids[i] < 0 || ids[i] >= s->N) return fail(PT_EINVAL, "initial infected id %d out of range", ids[i]);
    {
        std::vector<int32_t> sorted(ids, ids + n);
        std::sort(sorted.begin(), sorted.end());
        for (int i = 1; i < n; ++i)
            if (sorted[i] == sorted[i - 1]) return fail(PT_EINVAL, "duplicate initial infected id %d", sorted[i]);
    }
    HIP_TRY(hipSetDevice(s->device));
    const auto t0 = std::chrono::steady_clock::now();
    if (n > 0) {
        int32_t* d_ids = nullptr;
        if (dev_alloc(&d_ids, n)) return PT_ENOMEM;
        HIP_TRY(hipMemcpyAsync(d_ids, ids, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, s->stream));
        k_seed_infections<<<(n + 255) / 256, 256, 0, s->stream>>>(d_ids, n_p, n_s, s->st_p, s->st_s, s->dy_p, s->dy_s,
                                                                  s->gn_p, s->gn_s);
        hipError_t e = hipStreamSynchronize(s->stream);
        (void)hipFree(d_ids);
        if (e != hipSuccess) return fail(PT_EDEVICE, "seed infections: %s", hipGetErrorString(e));
    }
    s->seeded = true;
    const int rc = reset_day_state(s);
    s->setup_ms[PT_SETUP_INITIAL_INFECTED] = ms_between(t0, std::chrono::steady_clock::now());
    return rc;
}

int pt_sim_step(pt_sim* s, int32_t counts16_out[16], int32_t* infected_out) {
    if (!s) return fail(PT_EINVAL, "null handle");
    if (!s->setup_done) return fail(PT_ESTATE, "step before setup");
    HIP_TRY(hipSetDevice(s->device));
    DayState h;
    if (read_state(s, &h)) return PT_EDEVICE;
    if (h.done) return 1;
    const int day = h.day;
    if (enqueue_day(s, day)) return PT_EDEVICE;
    if (read_state(s, &h)) return PT_EDEVICE;
    s->timer.collect();
    if (h.fault) return fail(PT_EINVAL, "a kernel saw an errand destination outside the model's range");
    int32_t c[16], inf = 0;
    HIP_TRY(hipMemcpy(c, s->d_counts + 16 * day, sizeof c, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(&inf, s->d_inf + day, sizeof inf, hipMemcpyDeviceToHost));
    if (counts16_out) memcpy(counts16_out, c, sizeof c);
    if (infected_out) *infected_out = inf;
    return inf == 0 ? 1 : 0;
}

int pt_sim_run(pt_sim* s, int32_t* counts_out, int32_t* infected_out, int32_t* days_out) {
    if (!s) return fail(PT_EINVAL, "null handle");
    if (!s->setup_done) return fail(PT_ESTATE, "run before setup");
    HIP_TRY(hipSetDevice(s->device));
    DayState h;
    if (read_state(s, &h)) return PT_EDEVICE;
    const int first = h.day, first_rows = h.days_run;
    HIP_TRY(hipEventRecord(s->run_begin, s->stream));
    // Days go out in groups of kRunGroup.  After each group the loop's done
    // flag is copied back; the host then waits for the group before (so one
    // group stays queued and the GPU never idles) and stops enqueuing once
    // that one had ended the loop.  Days enqueued after the end are no-ops
    // (every kernel checks the flag), so at most two groups run idle.
    constexpr int kRunGroup = 4;
    if (!h.done) {
        int d = first;
        for (int g = 0; d < s->params.max_days; ++g) {
            const int end = std::min(d + kRunGroup, s->params.max_days);
            for (; d < end; ++d)
                if (enqueue_day(s, d)) return PT_EDEVICE;
            HIP_TRY(hipMemcpyAsync(&s->h_done[g & 1], &s->d_state->done, sizeof(int32_t), hipMemcpyDeviceToHost,
                                   s->stream));
            HIP_TRY(hipEventRecord(s->group_ev[g & 1], s->stream));
            if (g > 0) {
                HIP_TRY(hipEventSynchronize(s->group_ev[(g - 1) & 1]));
                if (s->h_done[(g - 1) & 1]) break;
            }
        }
    }
    HIP_TRY(hipEventRecord(s->run_end, s->stream));
    HIP_TRY(hipEventSynchronize(s->run_end));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, s->run_begin, s->run_end));
    s->last_run_ms = ms;
    s->timer.collect();
    if (read_state(s, &h)) return PT_EDEVICE;
    if (h.fault) return fail(PT_EINVAL, "a kernel saw an errand destination outside the model's range");
    const int rows = h.days_run;
    if (counts_out && rows > first_rows)
        HIP_TRY(hipMemcpy(counts_out, s->d_counts + 16 * first_rows, sizeof(int32_t) * 16 * (rows - first_rows),
                          hipMemcpyDeviceToHost));
    if (infected_out && rows > first_rows)
        HIP_TRY(hipMemcpy(infected_out, s->d_inf + first_rows, sizeof(int32_t) * (rows - first_rows),
                          hipMemcpyDeviceToHost));
    if (days_out) *days_out = rows - first_rows;
    return PT_OK;
}

int pt_sim_generation_offsets(pt_sim* s, int32_t* c_p, int32_t* c_s, int32_t len) {
    if (!s || !c_p || !c_s) return fail(PT_EINVAL, "null argument");
    if (!s->setup_done) return fail(PT_ESTATE, "generation offsets before setup");
    const int md = s->params.max_days;
    if (len < md + 1) return fail(PT_EINVAL, "len must be >= max_days + 1");
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipMemsetAsync(s->d_scratch64, 0, 512 * sizeof(unsigned long long), s->stream));
    k_gen_hist<<<grid_for(s->N, 256, 1024), 256, 0, s->stream>>>(s->gn_p, s->gn_s, s->N, s->d_scratch64);
    unsigned long long hist[512];
    HIP_TRY(hipMemcpyAsync(hist, s->d_scratch64, sizeof hist, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    int64_t ap = 0, as = 0;
    for (int k = 0; k < md; ++k) {
        c_p[k] = (int32_t)ap;
        c_s[k] = (int32_t)as;
        ap += (int64_t)hist[k];
        as += (int64_t)hist[256 + k];
    }
    c_p[md] = (int32_t)s->N;
    c_s[md] = (int32_t)s->N;
    return PT_OK;
}

int64_t pt_sim_array_bytes(pt_sim* s, int32_t which) {
    if (!s || !s->setup_done) return -1;
    const int64_t N = s->N;
    switch (which) {
        case PT_STATUS_P: case PT_STATUS_S: case PT_DAY_P: case PT_DAY_S: case PT_GEN_P: case PT_GEN_S:
        case PT_AGE: return N;
        case PT_HOUSEHOLD: return 4 * N;
        case PT_HH_OFFSETS: return 4 * ((int64_t)s->H + 1);
        case PT_WP_OFFSETS: return 4 * ((int64_t)s->W + 1);
        case PT_WP_PEOPLE: return 4 * N;
        case PT_INFECTED: {
            DayState h;
            if (read_state(s, &h)) return -1;
            return s->host_day > 0 ? 4 * (int64_t)h.n_infected : 0;
        }
        case PT_ERRAND_OFFSETS:
            if (s->last_weekend < 0) return 0;
            return 4 * ((int64_t)(s->last_weekend ? kWeHours : kWdHours) * s->W + 1);
        case PT_ERRAND_PEOPLE: return s->last_weekend < 0 ? 0 : 4 * s->last_table_n;
        case PT_HH_TYPES: return (int64_t)s->H + 1;
        case PT_WORKPLACE: return 4 * N;
        case PT_ERRANDS:
            if (s->last_weekend < 0) return 0;
            return 4 * N * (s->last_weekend ? 3 : 2);
        case PT_LOCATION_TYPE_OFFSETS: return 4 * (kTypes + 1);
    }
    return -1;
}

int pt_sim_copy_array(pt_sim* s, int32_t which, void* dst, int64_t bytes) {
    if (!s || !dst) return fail(PT_EINVAL, "null argument");
    if (!s->setup_done) return fail(PT_ESTATE, "copy before setup");
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    const int64_t need = pt_sim_array_bytes(s, which);
    if (need < 0) return fail(PT_EINVAL, "unknown array %d", which);
    if (bytes < need) return fail(PT_EINVAL, "destination holds %lld bytes, %lld needed", (long long)bytes,
                                  (long long)need);
    if (need == 0) return PT_OK;
    const void* src = nullptr;
    switch (which) {
        case PT_STATUS_P: src = s->st_p; break;
        case PT_STATUS_S: src = s->st_s; break;
        case PT_DAY_P: src = s->dy_p; break;
        case PT_DAY_S: src = s->dy_s; break;
        case PT_GEN_P: src = s->gn_p; break;
        case PT_GEN_S: src = s->gn_s; break;
        case PT_AGE: src = s->age; break;
        case PT_HOUSEHOLD: src = s->household; break;
        case PT_HH_OFFSETS: src = s->hh_off; break;
        case PT_WP_OFFSETS: src = s->wp_off; break;
        case PT_WP_PEOPLE: src = s->wp_people; break;
        case PT_INFECTED: src = s->infected; break;
        case PT_HH_TYPES: src = s->hh_types; break;
        case PT_WORKPLACE:
        case PT_ERRANDS: {
            // Computed into the key scratch (3N words, free between days).
            if (which == PT_WORKPLACE)
                k_workplace_of<<<grid_for(s->N, 256, 8192), 256, 0, s->stream>>>(s->d_model, s->workplace_off, s->age,
                                                                                s->N, (int32_t*)s->keys);
            else
                k_errands_of<<<grid_for(s->N, 256, 8192), 256, 0, s->stream>>>(
                    s->d_model, (const DayState*)s->d_state, s->last_weekend, s->age, s->N, (int32_t*)s->keys);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipStreamSynchronize(s->stream));
            src = s->keys;
            break;
        }
        case PT_LOCATION_TYPE_OFFSETS: {
            int32_t* out = (int32_t*)dst;
            for (int t = 0; t < kTypes; ++t) out[t] = s->model.wp_offset[t];
            out[kTypes] = s->W;
            return PT_OK;
        }
        case PT_ERRAND_PEOPLE:
            if (s->last_weekend == 1) {
                src = s->er_people;
                break;
            }
            return expand_weekday_table(s, nullptr, (int32_t*)dst, need / 4);
        case PT_ERRAND_OFFSETS: {
            if (s->last_weekend == 0) return expand_weekday_table(s, (int32_t*)dst, nullptr, 0);
            // Expand the dense bucket layout to the reference's
            // hour * W + location layout (lower_bound semantics).
            const int E = s->model.we_E, base = s->model.we_base;
            std::vector<int32_t> dense((size_t)kWeHours * E + 1);
            HIP_TRY(hipMemcpy(dense.data(), s->er_off, dense.size() * 4, hipMemcpyDeviceToHost));
            int32_t* out = (int32_t*)dst;
            for (int hr = 0; hr < kWeHours; ++hr)
                for (int l = 0; l < s->W; ++l)
                    out[(int64_t)hr * s->W + l] = l < base ? dense[(size_t)hr * E] : dense[(size_t)hr * E + (l - base)];
            out[(int64_t)kWeHours * s->W] = dense[(size_t)kWeHours * E];
            return PT_OK;
        }
        default: return fail(PT_EINVAL, "unknown array %d", which);
    }
    HIP_TRY(hipMemcpy(dst, src, (size_t)need, hipMemcpyDeviceToHost));
    return PT_OK;
}

int pt_sim_get_scalars(pt_sim* s, int64_t out[8]) {
    if (!s || !out) return fail(PT_EINVAL, "null argument");
    DayState h{};
    if (s->setup_done && read_state(s, &h)) return PT_EDEVICE;
    out[0] = s->N;
    out[1] = s->H;
    out[2] = s->W;
    out[3] = (int64_t)h.rand_offset;
    out[4] = (int64_t)s->workplace_off;
    out[5] = (int64_t)h.errand_off;
    out[6] = h.day;
    out[7] = h.days_run;
    return PT_OK;
}

int pt_sim_get_layout(pt_sim* s, int64_t out[8]) {
    if (!s || !out) return fail(PT_EINVAL, "null argument");
    const Model& m = s->model;
    out[0] = m.wd_base;
    out[1] = m.wd_E;
    out[2] = m.we_base;
    out[3] = m.we_E;
    out[4] = m.wd_K;
    out[5] = (int64_t)kWeHours * m.we_E;
    out[6] = radix_passes((int32_t)out[4], s->tables.max_bits);
    out[7] = radix_passes((int32_t)out[5], s->tables.max_bits);
    return PT_OK;
}

int pt_sim_snapshot(pt_sim* s) {
    if (!s) return fail(PT_EINVAL, "null handle");
    if (!s->setup_done) return fail(PT_ESTATE, "snapshot before setup");
    HIP_TRY(hipSetDevice(s->device));
    const int64_t Npad = ((s->N + 15) / 16) * 16;
    uint8_t* arr[6] = {s->st_p, s->st_s, s->dy_p, s->dy_s, s->gn_p, s->gn_s};
    for (int i = 0; i < 6; ++i)
        k_copy16<<<grid_for(Npad / 16, 256, 8192), 256, 0, s->stream>>>(
            reinterpret_cast<const uint4*>(arr[i]), reinterpret_cast<uint4*>(s->snap + i * Npad), Npad / 16);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->have_snapshot = true;
    return PT_OK;
}

int pt_sim_restore(pt_sim* s) {
    if (!s) return fail(PT_EINVAL, "null handle");
    if (!s->have_snapshot) return fail(PT_ESTATE, "restore without snapshot");
    HIP_TRY(hipSetDevice(s->device));
    const int64_t N = s->N, Npad = ((N + 15) / 16) * 16;
    uint8_t* arr[6] = {s->st_p, s->st_s, s->dy_p, s->dy_s, s->gn_p, s->gn_s};
    for (int i = 0; i < 6; ++i)
        k_copy16<<<grid_for(Npad / 16, 256, 8192), 256, 0, s->stream>>>(
            reinterpret_cast<const uint4*>(s->snap + i * Npad), reinterpret_cast<uint4*>(arr[i]), Npad / 16);
    k_fill_u64<<<grid_for(N, 256, 65535), 256, 0, s->stream>>>(s->claim_p, N, ~0ull);
    k_fill_u64<<<grid_for(N, 256, 65535), 256, 0, s->stream>>>(s->claim_s, N, ~0ull);
    HIP_TRY(hipGetLastError());
    return reset_day_state(s);
}

// ---------------- checkpoint / resume ----------------
//
// Buffer layout (host byte order, little-endian on every supported host):
//   CkptHeader
//   Model          the handle's model after setup (scales, tables, seed
//                  key, interventions, N): a checkpoint loads only into a
//                  handle built from the same inputs
//   DayState       the loop's device counters at a day boundary
//   int32 counts[256 * 16], infected[256], tab_full[256], tab_pairs[256]
//   uint8 status_p[N], status_s[N], day_p[N], day_s[N], gen_p[N], gen_s[N]
// The population (households, ages, workplaces) is not stored: it is a
// function of the inputs, rebuilt by pt_sim_setup.  Claim words are empty
// at every day boundary and are cleared on load.
namespace {

struct CkptHeader {
    char magic[8];
    uint32_t version;
    uint32_t model_bytes;
    uint32_t state_bytes;
    uint32_t rows_capacity;
    int64_t n_people;
    uint64_t checksum;  // over everything after the header
};
constexpr char kCkptMagic[8] = {'P', 'T', 'S', 'T', 'A', 'T', 'E', '1'};
constexpr uint32_t kCkptVersion = 1;
constexpr int kCkptRows = 256;
constexpr int64_t kCkptLogBytes = (int64_t)kCkptRows * (16 + 1 + 1 + 1) * 4;

int64_t ckpt_fixed_bytes() { return (int64_t)sizeof(CkptHeader) + sizeof(Model) + sizeof(DayState) + kCkptLogBytes; }

// FNV-1a over 64-bit words (then the tail bytes).
uint64_t ckpt_hash(const uint8_t* p, int64_t n) {
    uint64_t h = 0xcbf29ce484222325ull;
    int64_t i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        memcpy(&w, p + i, 8);
        h = (h ^ w) * 0x100000001b3ull;
    }
    for (; i < n; ++i) h = (h ^ p[i]) * 0x100000001b3ull;
    return h;
}

}  // namespace

int64_t pt_sim_checkpoint_bytes(pt_sim* s) {
    if (!s || !s->setup_done) return -1;
    return ckpt_fixed_bytes() + 6 * s->N;
}

int pt_sim_checkpoint_save(pt_sim* s, void* buf, int64_t bytes) {
    if (!s || !buf) return fail(PT_EINVAL, "null argument");
    if (!s->setup_done || !s->seeded) return fail(PT_ESTATE, "checkpoint before setup and seeding");
    const int64_t need = pt_sim_checkpoint_bytes(s);
    if (bytes < need) return fail(PT_EINVAL, "checkpoint buffer holds %lld bytes, %lld needed", (long long)bytes,
                                  (long long)need);
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    uint8_t* out = static_cast<uint8_t*>(buf);
    CkptHeader hd{};
    memcpy(hd.magic, kCkptMagic, 8);
    hd.version = kCkptVersion;
    hd.model_bytes = sizeof(Model);
    hd.state_bytes = sizeof(DayState);
    hd.rows_capacity = kCkptRows;
    hd.n_people = s->N;
    uint8_t* p = out + sizeof hd;
    memcpy(p, &s->model, sizeof(Model));
    p += sizeof(Model);
    HIP_TRY(hipMemcpy(p, s->d_state, sizeof(DayState), hipMemcpyDeviceToHost));
    p += sizeof(DayState);
    HIP_TRY(hipMemcpy(p, s->d_counts, 4 * kCkptRows * 16, hipMemcpyDeviceToHost));
    p += 4 * kCkptRows * 16;
    HIP_TRY(hipMemcpy(p, s->d_inf, 4 * kCkptRows, hipMemcpyDeviceToHost));
    p += 4 * kCkptRows;
    HIP_TRY(hipMemcpy(p, s->d_tab_full, 4 * kCkptRows, hipMemcpyDeviceToHost));
    p += 4 * kCkptRows;
    HIP_TRY(hipMemcpy(p, s->d_tab_pairs, 4 * kCkptRows, hipMemcpyDeviceToHost));
    p += 4 * kCkptRows;
    const uint8_t* arr[6] = {s->st_p, s->st_s, s->dy_p, s->dy_s, s->gn_p, s->gn_s};
    for (int i = 0; i < 6; ++i) {
        HIP_TRY(hipMemcpy(p, arr[i], (size_t)s->N, hipMemcpyDeviceToHost));
        p += s->N;
    }
    hd.checksum = ckpt_hash(out + sizeof hd, need - (int64_t)sizeof hd);
    memcpy(out, &hd, sizeof hd);
    return PT_OK;
}

int pt_sim_checkpoint_load(pt_sim* s, const void* buf, int64_t bytes) {
    if (!s || !buf) return fail(PT_EINVAL, "null argument");
    if (!s->setup_done) return fail(PT_ESTATE, "checkpoint load before setup");
    const uint8_t* in = static_cast<const uint8_t*>(buf);
    if (bytes < (int64_t)sizeof(CkptHeader)) return fail(PT_EINVAL, "checkpoint truncated (%lld bytes)", (long long)bytes);
    CkptHeader hd;
    memcpy(&hd, in, sizeof hd);
    if (memcmp(hd.magic, kCkptMagic, 8) != 0) return fail(PT_EINVAL, "not a checkpoint (bad magic)");
    if (hd.version != kCkptVersion || hd.model_bytes != sizeof(Model) || hd.state_bytes != sizeof(DayState) ||
        hd.rows_capacity != (uint32_t)kCkptRows)
        return fail(PT_EINVAL, "checkpoint written by an incompatible library version");
    if (hd.n_people != s->N)
        return fail(PT_EINVAL, "checkpoint holds %lld agents, this simulation %lld", (long long)hd.n_people,
                    (long long)s->N);
    const int64_t need = pt_sim_checkpoint_bytes(s);
    if (bytes < need) return fail(PT_EINVAL, "checkpoint truncated (%lld of %lld bytes)", (long long)bytes,
                                  (long long)need);
    if (ckpt_hash(in + sizeof hd, need - (int64_t)sizeof hd) != hd.checksum)
        return fail(PT_EINVAL, "checkpoint checksum mismatch (corrupt file)");
    const uint8_t* p = in + sizeof hd;
    if (memcmp(p, &s->model, sizeof(Model)) != 0)
        return fail(PT_EINVAL, "checkpoint was written for other constants, seed or interventions");
    p += sizeof(Model);
    DayState st;
    memcpy(&st, p, sizeof st);
    p += sizeof(DayState);
    if (st.day < 0 || st.day > 256 || st.days_run < 0 || st.days_run > kCkptRows || st.max_days != s->params.max_days)
        return fail(PT_EINVAL, "checkpoint day state out of range");
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(s->d_state, &st, sizeof st, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->d_counts, p, 4 * kCkptRows * 16, hipMemcpyHostToDevice));
    p += 4 * kCkptRows * 16;
    HIP_TRY(hipMemcpy(s->d_inf, p, 4 * kCkptRows, hipMemcpyHostToDevice));
    p += 4 * kCkptRows;
    HIP_TRY(hipMemcpy(s->d_tab_full, p, 4 * kCkptRows, hipMemcpyHostToDevice));
    p += 4 * kCkptRows;
    HIP_TRY(hipMemcpy(s->d_tab_pairs, p, 4 * kCkptRows, hipMemcpyHostToDevice));
    p += 4 * kCkptRows;
    uint8_t* arr[6] = {s->st_p, s->st_s, s->dy_p, s->dy_s, s->gn_p, s->gn_s};
    for (int i = 0; i < 6; ++i) {
        HIP_TRY(hipMemcpy(arr[i], p, (size_t)s->N, hipMemcpyHostToDevice));
        p += s->N;
    }
    k_fill_u64<<<grid_for(s->N, 256, 65535), 256, 0, s->stream>>>(s->claim_p, s->N, ~0ull);
    k_fill_u64<<<grid_for(s->N, 256, 65535), 256, 0, s->stream>>>(s->claim_s, s->N, ~0ull);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->seeded = true;
    s->host_day = 0;
    s->last_weekend = -1;
    return PT_OK;
}

int pt_sim_daily_rows(pt_sim* s, int32_t* counts_out, int32_t* infected_out, int32_t max_rows) {
    if (!s) return fail(PT_EINVAL, "null handle");
    if (!s->setup_done) return fail(PT_ESTATE, "daily rows before setup");
    HIP_TRY(hipSetDevice(s->device));
    DayState h;
    if (read_state(s, &h)) return PT_EDEVICE;
    const int rows = std::max(0, std::min(max_rows, h.days_run));
    if (rows > 0 && counts_out)
        HIP_TRY(hipMemcpy(counts_out, s->d_counts, sizeof(int32_t) * 16 * rows, hipMemcpyDeviceToHost));
    if (rows > 0 && infected_out)
        HIP_TRY(hipMemcpy(infected_out, s->d_inf, sizeof(int32_t) * rows, hipMemcpyDeviceToHost));
    return rows;
}

int pt_sim_setup_times(pt_sim* s, double* ms_out, int32_t n) {
    if (!s || !ms_out) return fail(PT_EINVAL, "null argument");
    for (int k = 0; k < n && k < PT_SETUP_PHASES; ++k) ms_out[k] = s->setup_ms[k];
    return PT_SETUP_PHASES;
}

int pt_sim_day_kernel_times(pt_sim* s, double* ms_out, int32_t max_days) {
    if (!s) return fail(PT_EINVAL, "null handle");
    const int days = (int)(s->timer.day_ms.size() / KC_COUNT);
    const int n = std::max(0, std::min(days, max_days));
    if (ms_out && n > 0) memcpy(ms_out, s->timer.day_ms.data(), sizeof(double) * KC_COUNT * n);
    return days;
}

int pt_device_memory(int32_t device, uint64_t* free_bytes, uint64_t* total_bytes) {
    HIP_TRY(hipSetDevice(device));
    size_t f = 0, t = 0;
    HIP_TRY(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return PT_OK;
}

int pt_device_name(int32_t device, char* buf, int32_t len) {
    if (!buf || len <= 0) return fail(PT_EINVAL, "null buffer");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    const char* name = prop.name[0] ? prop.name : prop.gcnArchName;
    snprintf(buf, (size_t)len, "%s", name);
    return PT_OK;
}

int pt_sim_set_interventions(pt_sim* s, const pt_interventions* iv) {
    if (!s || !iv) return fail(PT_EINVAL, "null argument");
    if (s->setup_done) return fail(PT_ESTATE, "interventions must be set before setup");
    Model m = s->model;
    const int rc = apply_interventions(iv, m);
    if (rc != PT_OK) return rc;
    s->model = m;
    return PT_OK;
}

int pt_sim_enable_trace(pt_sim* s, int32_t on) {
    if (!s) return fail(PT_EINVAL, "null handle");
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (!on) {
        free_trace(s);
        s->trace_on = false;
        return PT_OK;
    }
    if (s->setup_done && !s->tr_victim && alloc_trace(s)) return fail(PT_ENOMEM, "trace allocation failed");
    s->trace_on = true;
    return PT_OK;
}

int64_t pt_sim_trace_size(pt_sim* s) {
    if (!s || !s->setup_done || !s->trace_on || s->host_day == 0 || s->last_weekend < 0) return 0;
    DayState h;
    if (read_state(s, &h)) return -1;
    return (int64_t)h.n_infected * (s->last_weekend ? 5 : 8);
}

int pt_sim_trace(pt_sim* s, int32_t* infector, int32_t* victim, uint8_t* type, int32_t* loc, uint8_t* action,
                 int64_t n) {
    if (!s) return fail(PT_EINVAL, "null handle");
    const int64_t have = pt_sim_trace_size(s);
    if (have < 0) return PT_EDEVICE;
    if (n < have) return fail(PT_EINVAL, "trace holds %lld entries, room for %lld", (long long)have, (long long)n);
    if (have == 0) return 0;
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    const int C = s->last_weekend ? 5 : 8;
    if (infector) {
        std::vector<int32_t> inf((size_t)(have / C));
        HIP_TRY(hipMemcpy(inf.data(), s->infected, inf.size() * 4, hipMemcpyDeviceToHost));
        for (int64_t e = 0; e < have; ++e) infector[e] = inf[(size_t)(e / C)];
    }
    if (victim) HIP_TRY(hipMemcpy(victim, s->tr_victim, (size_t)have * 4, hipMemcpyDeviceToHost));
    if (type) HIP_TRY(hipMemcpy(type, s->tr_type, (size_t)have, hipMemcpyDeviceToHost));
    if (loc) HIP_TRY(hipMemcpy(loc, s->tr_loc, (size_t)have * 4, hipMemcpyDeviceToHost));
    if (action) HIP_TRY(hipMemcpy(action, s->tr_action, (size_t)have, hipMemcpyDeviceToHost));
    return (int)std::min<int64_t>(have, 0x7FFFFFFF);
}

int pt_sim_trace_draws(pt_sim* s, uint8_t* draw, uint8_t* profile, int64_t n) {
    if (!s) return fail(PT_EINVAL, "null handle");
    const int64_t have = pt_sim_trace_size(s);
    if (have < 0) return PT_EDEVICE;
    if (n < have) return fail(PT_EINVAL, "trace holds %lld entries, room for %lld", (long long)have, (long long)n);
    if (have == 0) return 0;
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (draw) HIP_TRY(hipMemcpy(draw, s->tr_draw, (size_t)have, hipMemcpyDeviceToHost));
    if (profile) HIP_TRY(hipMemcpy(profile, s->tr_prof, (size_t)have, hipMemcpyDeviceToHost));
    return (int)std::min<int64_t>(have, 0x7FFFFFFF);
}

int pt_sim_set_launch(pt_sim* s, int32_t knob, int32_t value) {
    if (!s) return fail(PT_EINVAL, "null handle");
    if (int rc = check_launch(knob, value)) return rc;
    s->launch[knob] = value;
    s->grid_contacts = s->cus * s->launch[PT_LAUNCH_CONTACT_BLOCKS_PER_CU];
    s->grid_gen = s->cus * s->launch[PT_LAUNCH_GEN_BLOCKS_PER_CU];
    s->tables.wide_bits = s->launch[PT_LAUNCH_SORT_WIDE_BITS];
    s->tables.max_bits = s->launch[PT_LAUNCH_SORT_MAX_BITS];
    return PT_OK;
}

int pt_sim_get_launch(pt_sim* s, int32_t knob) {
    if (!s) return fail(PT_EINVAL, "null handle");
    if (knob < 0 || knob >= PT_LAUNCH_KNOBS) return fail(PT_EINVAL, "launch setting %d unknown", knob);
    return s->launch[knob];
}

int pt_sim_set_table_mode(pt_sim* s, int32_t mode) {
    if (!s) return fail(PT_EINVAL, "null handle");
    if (mode < TABLES_AUTO || mode > TABLES_SUBSET) return fail(PT_EINVAL, "table mode %d unknown", mode);
    s->table_mode = mode;
    return PT_OK;
}

int pt_sim_table_log(pt_sim* s, int32_t* full_out, int32_t* pairs_out, int32_t len) {
    if (!s) return fail(PT_EINVAL, "null handle");
    if (!s->setup_done) return fail(PT_ESTATE, "table log before setup");
    HIP_TRY(hipSetDevice(s->device));
    DayState h;
    if (read_state(s, &h)) return PT_EDEVICE;
    // The terminal I == 0 day builds no table and is not logged.
    const int days = std::max(0, std::min(len, h.done && h.n_infected == 0 ? h.days_run - 1 : h.days_run));
    if (days > 0 && full_out) HIP_TRY(hipMemcpy(full_out, s->d_tab_full, sizeof(int32_t) * days, hipMemcpyDeviceToHost));
    if (days > 0 && pairs_out)
        HIP_TRY(hipMemcpy(pairs_out, s->d_tab_pairs, sizeof(int32_t) * days, hipMemcpyDeviceToHost));
    return days;
}

int pt_sim_enable_timing(pt_sim* s, int32_t enable) {
    if (!s) return fail(PT_EINVAL, "null handle");
    s->timer.reset();
    s->timer.mode = enable < 0 ? -1 : enable;
    return PT_OK;
}

int pt_sim_kernel_times(pt_sim* s, double* ms_out, int64_t* launches_out, int32_t n) {
    if (!s) return fail(PT_EINVAL, "null handle");
    for (int k = 0; k < n && k < KC_COUNT; ++k) {
        if (ms_out) ms_out[k] = s->timer.ms[k];
        if (launches_out) launches_out[k] = s->timer.launches[k];
    }
    return KC_COUNT;
}

const char* pt_sim_kernel_name(int32_t k) { return (k >= 0 && k < KC_COUNT) ? kKernelNames[k] : ""; }

int pt_sim_synchronize(pt_sim* s) {
    if (!s) return fail(PT_EINVAL, "null handle");
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return PT_OK;
}

double pt_sim_last_run_ms(pt_sim* s) { return s ? s->last_run_ms : -1.0; }

void pt_sim_destroy(pt_sim* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    free_population(s);
    dev_free(s->d_model);
    dev_free(s->d_state);
    dev_free(s->d_counts);
    dev_free(s->d_inf);
    dev_free(s->d_tab_full);
    dev_free(s->d_tab_pairs);
    dev_free(s->d_scratch64);
    if (s->run_begin) (void)hipEventDestroy(s->run_begin);
    if (s->run_end) (void)hipEventDestroy(s->run_end);
    for (auto e : s->setup_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto e : s->group_ev)
        if (e) (void)hipEventDestroy(e);
    if (s->h_done) (void)hipHostFree(s->h_done);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

}  // extern "C"

