// pt_day.h — daily-loop device code shared by the single-simulation engine
// (pt_sim.hip) and the replicate-batch engine (pt_batch.hip), and the host
// helpers both use.
//
// Every function here works on one simulation in its own (local) agent,
// household, workplace and bucket numbering; the batch engine reaches a
// replicate by offsetting the array pointers and passing a bucket or id
// bias.
//
// Citations are relative to PandemicThrust/PandemicThrust/ of the reference.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/pandemic/pt_abi.h"
#include "pt_model.h"
#include "pt_sort.h"

namespace pt {

// ---------------- host helpers ----------------

// Records the error text returned by pt_last_error; returns code.
int fail(int code, const char* fmt, ...);
// PT_OK when value is in range for launch setting knob (pt_launch_knob).
int check_launch(int32_t knob, int32_t value);

// Argument checks shared by the location-table entry points (pt_table.hip).
int location_table_args(const uint32_t* keys, int64_t n, int32_t n_buckets, const int32_t* people,
                        const int32_t* offsets);

// Makes `device` current for a scope and restores the caller's device.
struct DeviceScope {
    int saved = -1;
    hipError_t err = hipSuccess;
    explicit DeviceScope(int device) {
        err = hipGetDevice(&saved);
        if (err == hipSuccess) err = hipSetDevice(device);
    }
    ~DeviceScope() {
        if (saved >= 0) (void)hipSetDevice(saved);
    }
    DeviceScope(const DeviceScope&) = delete;
    DeviceScope& operator=(const DeviceScope&) = delete;
};

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return ::pt::fail(PT_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

template <class T>
int dev_alloc(T** p, int64_t count) {
    if (count <= 0) count = 1;
    if (hipMalloc((void**)p, (size_t)count * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        return fail(PT_ENOMEM, "hipMalloc of %lld bytes failed", (long long)(count * (int64_t)sizeof(T)));
    }
    return 0;
}

template <class T>
void dev_free(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

inline int grid_for(int64_t n, int threads, int cap) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + threads - 1) / threads, cap));
}

// Model tables of one constants row (setup_loadParameters PandemicSim.cu:313-562,
// setup_scaleSimulation :1127-1162, setup_calculateInfectionData :2828-2840)
// and the dense errand-table layout.  Returns PT_OK or PT_EINVAL.
int build_model(const pt_params* p, const int32_t seed_key[2], Model& m);

// Validates an intervention set and writes it into the model (PT_OK or
// PT_EINVAL).
int apply_interventions(const pt_interventions* iv, Model& m);

// Population arrays of one simulation (local numbering).
struct PopArrays {
    uint8_t* hh_types;   // H + 1
    int32_t* hh_off;     // H + 1
    int32_t* household;  // N
    uint8_t* age;        // N
    int32_t* wp_off;     // W + 1
    int32_t* wp_people;  // N
    uint8_t *st_p, *st_s, *dy_p, *dy_s, *gn_p, *gn_s;  // N
    uint64_t *claim_p, *claim_s;                        // N
};

// K1/K2 (PandemicSim.cu:3579-3605, :2329-2357): household types into
// hh_types and the population size added to *people (device counter).
void launch_household_types(const Model* d_model, int H, uint8_t* hh_types, unsigned long long* people,
                            hipStream_t stream);

// K3-K7 (PandemicSim.cu:2573-2613, :3812-3862, :3654-3664) for a model whose
// N is set on the device: household offsets and people, workplaces and the
// workplace table, statuses, empty claim words.  tmp holds max(N, H + 1)
// uint32; ts is sized for N pairs.  Enqueued on stream, not synchronised.
// marks, if given, are 4 events recorded before the households, after them,
// after the workplaces and after the statuses.
hipError_t build_population(const Model* d_model, int64_t N, int H, int W, const PopArrays& P, uint32_t* tmp,
                            TableScratch& ts, uint64_t workplace_off, hipStream_t stream,
                            const hipEvent_t* marks = nullptr);

// ---------------- device state ----------------

constexpr int kSweepThreads = 512;
constexpr int kSweepPer = 16;                          // agents per thread (one uint4 per array)
constexpr int kSweepTile = kSweepThreads * kSweepPer;  // 8192 agents per block
constexpr int kPlanThreads = 1024;
constexpr int kContactThreads = 256;

// The reference's host-side loop counters (PandemicSim.h:73-76,
// simRandOffsetsStruct_t) kept in device memory.
struct DayState {
    int32_t day;           // current_day
    int32_t done;          // the loop has ended
    int32_t days_run;      // rows of output_infected_stats.csv produced
    int32_t n_infected;    // infected_count of the current day
    uint64_t rand_offset;  // rand_offset
    uint64_t errand_off;   // errand_randOffset of the current day
    uint64_t contact_base; // rand_offset at the contact kernel
    uint32_t n_claimed[2]; // claim-list lengths (pandemic, seasonal)
    int32_t max_days;
    int32_t fault;         // nonzero if a kernel saw data outside the model's range
    int32_t skip_full;     // the full errand-table path is idle today
    int32_t skip_subset;   // the subset errand-table path is idle today
    uint32_t n_sel;        // pairs selected into today's subset table
    uint32_t n_marked;     // errand buckets queried today
    int32_t skip_mark;     // too many infectors today for a subset table to pay off
};

// Errand tables are built either for every bucket (full) or only for the
// buckets today's infectors draw contacts from (subset); the choice is
// made on the device from the number of queried buckets.
enum TableMode : int { TABLES_AUTO = 0, TABLES_FULL = 1, TABLES_SUBSET = 2 };
constexpr int kSelThreads = 256;
constexpr int kSelPerWd = 8;   // agents per thread, weekday selection
constexpr int kSelPerWe = 4;   // agents per thread, weekend selection
constexpr int kSelAgentsWd = kSelThreads * kSelPerWd;
constexpr int kSelAgentsWe = kSelThreads * kSelPerWe;
constexpr int kSelStageWd = kSelAgentsWd * 2;  // staged pairs per weekday selection block
constexpr int kSelStageWe = kSelAgentsWe * 3;  // staged pairs per weekend selection block

// Whether a subset table pays off: fewer than 60% of the buckets queried.
__host__ __device__ __forceinline__ bool subset_pays(int64_t marked, int64_t K, int share_pct) {
    return marked * 100 < K * share_pct;
}

// Infected count up to which a day marks its queried buckets (K * pct / 100).
__host__ __device__ __forceinline__ int64_t mark_limit(int64_t K, int pct) { return K * pct / 100; }

// ---------------- wave helpers ----------------

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Sums over the 64 lanes of eight byte counters packed in v (bin b in byte
// b; at most 16 per lane): three butterfly steps on the packed words (a byte
// reaches at most 128), then three on 16-bit fields.  out[b] = the total of
// bin b, in every lane.
__device__ __forceinline__ void wave_sum_bins8(uint64_t v, uint32_t (&out)[8]) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
#pragma unroll
    for (int d = 1; d < 8; d <<= 1) {
        lo += __shfl_xor(lo, d, 64);
        hi += __shfl_xor(hi, d, 64);
    }
    uint32_t w[4] = {lo & 0x00FF00FFu, (lo >> 8) & 0x00FF00FFu, hi & 0x00FF00FFu, (hi >> 8) & 0x00FF00FFu};
#pragma unroll
    for (int d = 8; d < 64; d <<= 1) {
#pragma unroll
        for (int j = 0; j < 4; ++j) w[j] += __shfl_xor(w[j], d, 64);
    }
    // w[0]: bins 0 and 2, w[1]: 1 and 3, w[2]: 4 and 6, w[3]: 5 and 7
    out[0] = w[0] & 0xFFFFu;
    out[2] = w[0] >> 16;
    out[1] = w[1] & 0xFFFFu;
    out[3] = w[1] >> 16;
    out[4] = w[2] & 0xFFFFu;
    out[6] = w[2] >> 16;
    out[5] = w[3] & 0xFFFFu;
    out[7] = w[3] >> 16;
}

// All 64 lanes must call: reserves n_mine slots per lane in a list whose
// length lives at *counter; returns the lane's first slot.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, uint32_t n_mine, int lane) {
    const uint32_t inc = wave_incl_scan(n_mine, lane);
    const uint32_t total = __shfl(inc, 63, 64);
    uint32_t base = 0;
    if (lane == 0 && total) base = atomicAdd(counter, total);
    base = __shfl(base, 0, 64);
    return base + inc - n_mine;
}

// ---------------- D1 + D2: sweep ----------------

// One sweep block (kSweepThreads threads) over agents [t0, t0 + kSweepTile)
// of a population of N: recover (infection day <= current_day - (latent +
// infectious days), the reference's current_day - 10), count the 8 status
// codes per strain and stage the ids (+ id_bias) of infected agents.
// (kernel_countInfectedStatusAndRecover PandemicSim.cu:2097-2192,
//  daily_buildInfectedArray_global :2207-2224.)  The status arrays hold
// whole uint4 words past N.
__device__ __forceinline__ void sweep_block(const Model& M, int64_t N, int day, int64_t t0, uint8_t* __restrict__ st_p,
                                            uint8_t* __restrict__ st_s, const uint8_t* __restrict__ dy_p,
                                            const uint8_t* __restrict__ dy_s, uint32_t* __restrict__ counts16,
                                            uint32_t* __restrict__ ninf_out, int32_t* __restrict__ stage,
                                            int32_t id_bias) {
    const int dtr_p = day - (M.lat[0] + M.dur[0]);
    const int dtr_s = day - (M.lat[1] + M.dur[1]);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t a0 = t0 + (int64_t)tid * kSweepPer;
    uint8_t sp[kSweepPer], ss[kSweepPer];
    const bool full = a0 + kSweepPer <= N;
    if (full) {
        const uint4 vp = *reinterpret_cast<const uint4*>(st_p + a0);
        const uint4 vs = *reinterpret_cast<const uint4*>(st_s + a0);
        memcpy(sp, &vp, 16);
        memcpy(ss, &vs, 16);
    } else {
#pragma unroll
        for (int i = 0; i < kSweepPer; ++i) {
            sp[i] = (a0 + i < N) ? st_p[a0 + i] : kSusceptible;
            ss[i] = (a0 + i < N) ? st_s[a0 + i] : kSusceptible;
        }
    }
    if (dtr_p >= 0 || dtr_s >= 0) {
        bool anyp = false, anys = false;
#pragma unroll
        for (int i = 0; i < kSweepPer; ++i) {
            anyp |= sp[i] >= kInfected;
            anys |= ss[i] >= kInfected;
        }
        if (anyp && dtr_p >= 0) {
            uint8_t dp[kSweepPer];
            if (full) {
                const uint4 v = *reinterpret_cast<const uint4*>(dy_p + a0);
                memcpy(dp, &v, 16);
            } else {
#pragma unroll
                for (int i = 0; i < kSweepPer; ++i) dp[i] = (a0 + i < N) ? dy_p[a0 + i] : kDayNone;
            }
            bool dirty = false;
#pragma unroll
            for (int i = 0; i < kSweepPer; ++i)
                if (sp[i] >= kInfected && (int)dp[i] <= dtr_p) {
                    sp[i] = kRecovered;
                    dirty = true;
                }
            if (dirty) {
                if (full) {
                    uint4 v;
                    memcpy(&v, sp, 16);
                    *reinterpret_cast<uint4*>(st_p + a0) = v;
                } else {
                    for (int i = 0; i < kSweepPer; ++i)
                        if (a0 + i < N) st_p[a0 + i] = sp[i];
                }
            }
        }
        if (anys && dtr_s >= 0) {
            uint8_t ds[kSweepPer];
            if (full) {
                const uint4 v = *reinterpret_cast<const uint4*>(dy_s + a0);
                memcpy(ds, &v, 16);
            } else {
#pragma unroll
                for (int i = 0; i < kSweepPer; ++i) ds[i] = (a0 + i < N) ? dy_s[a0 + i] : kDayNone;
            }
            bool dirty = false;
#pragma unroll
            for (int i = 0; i < kSweepPer; ++i)
                if (ss[i] >= kInfected && (int)ds[i] <= dtr_s) {
                    ss[i] = kRecovered;
                    dirty = true;
                }
            if (dirty) {
                if (full) {
                    uint4 v;
                    memcpy(&v, ss, 16);
                    *reinterpret_cast<uint4*>(st_s + a0) = v;
                } else {
                    for (int i = 0; i < kSweepPer; ++i)
                        if (a0 + i < N) st_s[a0 + i] = ss[i];
                }
            }
        }
    }
    // Counts: 8 bins per strain packed as bytes (at most 16 per thread).
    uint64_t cp = 0, cs = 0;
    uint32_t inf_mask = 0;
#pragma unroll
    for (int i = 0; i < kSweepPer; ++i) {
        if (a0 + i < N) {
            cp += 1ull << (8 * (sp[i] & 7));
            cs += 1ull << (8 * (ss[i] & 7));
            if (sp[i] >= kInfected || ss[i] >= kInfected) inf_mask |= 1u << i;
        }
    }
    __shared__ uint32_t s_cnt[kSweepThreads / 64][16];
    __shared__ uint32_t s_inf[kSweepThreads / 64];
    {
        uint32_t vp[8], vs[8];
        wave_sum_bins8(cp, vp);
        wave_sum_bins8(cs, vs);
        if (lane == 0) {
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                s_cnt[wave][b] = vp[b];
                s_cnt[wave][8 + b] = vs[b];
            }
        }
    }
    const uint32_t ninf = (uint32_t)__popc(inf_mask);
    const uint32_t inc = wave_incl_scan(ninf, lane);
    if (lane == 63) s_inf[wave] = inc;
    __syncthreads();
    uint32_t wbase = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kSweepThreads / 64; ++w) {
        if (w < wave) wbase += s_inf[w];
        total += s_inf[w];
    }
    uint32_t pos = wbase + inc - ninf;
#pragma unroll
    for (int i = 0; i < kSweepPer; ++i)
        if (inf_mask & (1u << i)) stage[pos++] = id_bias + (int32_t)(a0 + i);
    if (tid < 16) {
        uint32_t s = 0;
#pragma unroll
        for (int w = 0; w < kSweepThreads / 64; ++w) s += s_cnt[w][tid];
        counts16[tid] = s;
    }
    if (tid == 0) *ninf_out = total;
}

// ---------------- errand-table selection ----------------

__device__ __forceinline__ bool is_marked(const uint32_t* __restrict__ mark, uint32_t key) {
    return (mark[key >> 5] >> (key & 31)) & 1u;
}

// One selection block (kSelThreads threads) over the agents from blk_a0 of a
// population of N: regenerates each agent's errand keys (weekday_keys or
// hour * E + location - base) and writes to sk/si, in ascending agent order,
// the pairs (kbias + key, agent) whose bucket kbias + key is marked, or every
// pair when all is set (through LDS, so the global writes are contiguous).
// Block-ordered output keeps the ids ascending within every bucket (equal
// buckets never come from two hours of one agent).  *cnt_out (if set) gets
// the number written.  Returns false if a destination fell outside the
// layout.
template <int PER, bool WEEKEND>
__device__ __forceinline__ bool select_block(const Model& M, uint64_t eoff, const uint8_t* __restrict__ age,
                                             const uint32_t* __restrict__ mark, uint32_t kbias, bool all,
                                             int64_t blk_a0, int64_t N, uint32_t* __restrict__ sk,
                                             int32_t* __restrict__ si, uint32_t* __restrict__ cnt_out) {
    constexpr int KPA = WEEKEND ? 3 : 2;  // keys per agent
    constexpr int NK = PER * KPA;
    const int64_t a0 = blk_a0 + (int64_t)threadIdx.x * PER;
    uint32_t key[NK];
    uint32_t sel = 0;
    bool bad = false;
    if (WEEKEND) {
        const int base = M.we_base, E = M.we_E;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int64_t p = a0 + q;
            key[3 * q] = key[3 * q + 1] = key[3 * q + 2] = 0;
            if (p < N) {
                int32_t hour[3], loc[3];
                weekend_errands(M, eoff + (uint64_t)(int64_t)(2 * (int32_t)p), hour, loc);
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    int32_t l = loc[e] - base;
                    if ((uint32_t)l >= (uint32_t)E) {
                        bad = true;
                        l = min(max(l, 0), E - 1);
                    }
                    const uint32_t k = kbias + (uint32_t)(hour[e] * E + l);
                    key[3 * q + e] = k;
                    if (all || is_marked(mark, k)) sel |= 1u << (3 * q + e);
                }
            }
        }
    } else {
        // The thread's ages in one load when its agents are all inside the
        // population (a0 is a multiple of PER; age is PER-byte aligned).
        uint8_t ag[PER];
        if (PER == 8 && a0 + PER <= N) {
            const uint2 v = *reinterpret_cast<const uint2*>(age + a0);
            memcpy(ag, &v, 8);
        } else {
#pragma unroll
            for (int q = 0; q < PER; ++q) ag[q] = a0 + q < N ? age[a0 + q] : 0;
        }
#pragma unroll
        for (int q = 0; q < PER / 2; ++q) {
            const int64_t g = a0 / 2 + q;
            key[4 * q] = key[4 * q + 1] = key[4 * q + 2] = key[4 * q + 3] = 0;
            if (2 * g < N) {
                const U4 u = draw4(eoff + (uint64_t)g, M.key0, M.key1);
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const int64_t p = 2 * g + k;
                    if (p < N) {
                        int32_t d1, d2;
                        weekday_dests(M, u.w[2 * k], u.w[2 * k + 1], ag[2 * q + k] == kAgeAdult, d1, d2);
                        const int i = 4 * q + 2 * k;
                        bad |= !weekday_keys(M, d1, d2, key[i], key[i + 1]);
                        key[i] += kbias;
                        key[i + 1] += kbias;
                        if (all || is_marked(mark, key[i])) sel |= 1u << i;
                        if (all || is_marked(mark, key[i + 1])) sel |= 1u << (i + 1);
                    }
                }
            }
        }
    }
    // Block-wide exclusive scan of the per-thread counts.
    __shared__ uint32_t s_w[kSelThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t cnt = (uint32_t)__popc(sel);
    const uint32_t inc = wave_incl_scan(cnt, lane);
    if (lane == 63) s_w[wave] = inc;
    __syncthreads();
    uint32_t pos = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kSelThreads / 64; ++w) {
        if (w < wave) pos += s_w[w];
        total += s_w[w];
    }
    pos += inc - cnt;
    // Staged at j + j / 16: dense blocks put thread t's pairs at 16t + i,
    // which the padding spreads over distinct LDS banks.
    constexpr int SN = kSelThreads * NK + kSelThreads * NK / 16;
    __shared__ uint32_t s_k[SN];
    __shared__ int32_t s_i[SN];
#pragma unroll
    for (int i = 0; i < NK; ++i)
        if (sel & (1u << i)) {
            s_k[pos + (pos >> 4)] = key[i];
            s_i[pos + (pos >> 4)] = (int32_t)(a0 + i / KPA);
            pos++;
        }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < total; j += kSelThreads) {
        sk[j] = s_k[j + (j >> 4)];
        si[j] = s_i[j + (j >> 4)];
    }
    if (cnt_out && threadIdx.x == 0) *cnt_out = total;
    return !bad;
}

// ---------------- contacts ----------------

// device_selectRandomPersonFromLocation (PandemicSim.cu:2028-2071) with
// device_getVictimAtIndex (:2016-2025): household members are consecutive ids.
__device__ __forceinline__ int select_person(int32_t infector, int32_t loc_off, int32_t loc_count, uint32_t r,
                                             int type, const int32_t* __restrict__ people, int32_t& victim) {
    if (loc_count > 1) {
        int vo = (int)(r % (uint32_t)loc_count);
        int32_t v = (type == CT_HOME) ? loc_off + vo : people[loc_off + vo];
        if (v == infector) {
            vo = (vo + 1) % loc_count;
            v = (type == CT_HOME) ? loc_off + vo : people[loc_off + vo];
        }
        victim = v;
        return type;
    }
    victim = loc_count;
    return CT_NONE;
}

// Where the victims of an infector's contact slots come from: the household
// (slots 0..2; its members are consecutive ids), then `wpd` workplace or
// school slots, then h0 slots in the first errand location and h1 in the
// second (weekend: wpd = 0, h0 = h1 = 1), with each slot's draw r.  Only a
// slot whose infection draw succeeds needs its victim, so outside
// validation mode the contact kernels look victims up then (pick) instead of
// gathering one from a location table for every slot: the contact types and
// kval_sum depend on the location counts alone (select_person), and the
// victims of unsuccessful slots are never read.
struct SlotPlan {
    int32_t hoff, hcnt, woff, wcnt, e0off, e0cnt, e1off, e1cnt;
    int wpd, h0, h1;
    uint32_t r[8];
};

// The victim of slot c (select_person's pick; the slot's count is > 1).
__device__ __forceinline__ int32_t pick(const SlotPlan& P, const int32_t* __restrict__ wp_people,
                                        const int32_t* __restrict__ er_people, int32_t infector, int c) {
    int32_t off, cnt;
    const int32_t* people;
    if (c < 3) {
        off = P.hoff;
        cnt = P.hcnt;
        people = nullptr;
    } else if (c < 3 + P.wpd) {
        off = P.woff;
        cnt = P.wcnt;
        people = wp_people;
    } else if (c < 3 + P.wpd + P.h0) {
        off = P.e0off;
        cnt = P.e0cnt;
        people = er_people;
    } else {
        off = P.e1off;
        cnt = P.e1cnt;
        people = er_people;
    }
    int32_t v;
    select_person(infector, off, cnt, P.r[c], people ? CT_ERRAND : CT_HOME, people, v);
    return v;
}

// The arrays of one simulation in its local numbering.  er_off is indexed by
// the simulation's bucket key; its values (and the other offsets) index the
// corresponding people arrays as given.
struct AgentArrays {
    const uint8_t* age;
    const int32_t* household;
    const int32_t* hh_off;
    const int32_t* wp_off;
    const int32_t* wp_people;
    const int32_t* er_off;
    const int32_t* er_people;
    const uint8_t* st_p;
    const uint8_t* st_s;
    const uint8_t* dy_p;
    const uint8_t* dy_s;
    const uint8_t* gn_p;
    const uint8_t* gn_s;
    uint64_t* claim_p;
    uint64_t* claim_s;
    // Optional per-contact trace of the day (validation mode), [I * C]
    // with C = 8 weekday / 5 weekend slots per infector: victim, type, the
    // contact's location (household, workplace, location id for
    // after-school, hour * W + location for errands; also for a contact that
    // found nobody else there, whose victim is the location's count; -1 for
    // an unused or dropped slot) and the
    // attempts that succeeded on a susceptible victim (2 pandemic, 1
    // seasonal; which of them won the victim is settled after all claims).
    int32_t* trace_victim;
    uint8_t* trace_type;
    int32_t* trace_loc;
    uint8_t* trace_attempt;
    // Optional per-slot draw record (validation mode), [I * C]: trace_draw
    // bit 1 / bit 0 = the pandemic / seasonal draw y < p * kval succeeded,
    // whatever the victim's state; trace_prof = the profile each strain's
    // draw would give (device_getInfectionProfile, PandemicSim.cu:3974-3988),
    // pandemic in the low nibble, seasonal in the high nibble.  The
    // reference computes both for every slot (PandemicSim.cu:3195-3211).
    uint8_t* trace_draw;
    uint8_t* trace_prof;
};

// Strain state of an infector (device_processContacts, PandemicSim.cu:3125-3175):
// profile and shedding day if infectious today, else -1.  Infectious while
// lat <= day - infection day < lat + dur (the reference: 0 <= . < 10), at
// shedding index day - infection day - lat.
struct Infectivity {
    int prof_p, dd_p, prof_s, dd_s;
    uint8_t gen_p, gen_s;
};

__device__ __forceinline__ Infectivity infectivity(const Model& M, const AgentArrays& A, int32_t idx, int day) {
    Infectivity f;
    f.prof_p = f.prof_s = -1;
    f.dd_p = f.dd_s = 0;
    f.gen_p = f.gen_s = kGenNone;
    const uint8_t sp = A.st_p[idx], ss = A.st_s[idx];
    if (sp >= kInfected) {
        const int dd = day - (int)A.dy_p[idx] - M.lat[0];
        if (dd >= 0 && dd < M.dur[0]) {
            f.prof_p = sp - kInfected;
            f.dd_p = dd;
            f.gen_p = (uint8_t)(A.gn_p[idx] + 1);
        }
    }
    if (ss >= kInfected) {
        const int dd = day - (int)A.dy_s[idx] - M.lat[1];
        if (dd >= 0 && dd < M.dur[1]) {
            f.prof_s = ss - kInfected;
            f.dd_s = dd;
            f.gen_s = (uint8_t)(A.gn_s[idx] + 1);
        }
    }
    return f;
}

// Whether the infector can infect at all today: with alpha*factor <= 0 every
// draw y >= 0 fails y < p*kval, so contacts need not be generated.
__device__ __forceinline__ bool can_infect(const Model& M, const Infectivity& f) {
    const bool p = f.prof_p >= 0 && M.shed[f.prof_p][f.dd_p] * M.factor[0] > 0.f;
    const bool s = f.prof_s >= 0 && M.shed[f.prof_s][f.dd_s] * M.factor[1] > 0.f;
    return p || s;
}

// Per-contact draws and claims (device_processContacts PandemicSim.cu:3177-3231).
// Every successful attempt on a victim that is susceptible at the start of
// the day posts ((pos*8 + slot) << 32 | gen << 8 | profile) with a 64-bit
// atomicMin into the victim's claim word; the first poster is told so.
// vic: the slots' victims when `resolved` (validation mode), else filled
// here for the slots whose draw succeeds (pick).
template <int C>
__device__ __forceinline__ void process_contacts(const Model& M, const AgentArrays& A, const Infectivity& f, int kval_sum,
                                                 int32_t (&vic)[8], const int (&typ)[8], const SlotPlan& P,
                                                 bool resolved, int32_t idx, uint64_t M2, uint32_t pos,
                                                 bool (&newp)[8], bool (&news)[8]) {
    const float prob_p = f.prof_p >= 0 ? infection_probability(M, f.prof_p, f.dd_p, 0, kval_sum) : -1.f;
    const float prob_s = f.prof_s >= 0 ? infection_probability(M, f.prof_s, f.dd_s, 1, kval_sum) : -1.f;
    uint32_t u[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const U4 w = draw4(M2 + (uint64_t)q, M.key0, M.key1);
#pragma unroll
        for (int j = 0; j < 4; ++j) u[4 * q + j] = w.w[j];
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const float kv = typ[c] == CT_NONE ? 0.f : 1.f;
        uint8_t attempt = 0;
        const bool draw_p = unit_float(u[2 * c]) < prob_p * kv;
        const bool draw_s = unit_float(u[2 * c + 1]) < prob_s * kv;
        int32_t v = vic[c];
        if (!resolved && (draw_p || draw_s)) {  // kv = 1: a real contact, count > 1
            v = pick(P, A.wp_people, A.er_people, idx, c);
            vic[c] = v;
        }
        if (A.trace_draw) {
            const int64_t e = (int64_t)pos * C + c;
            A.trace_draw[e] = (uint8_t)((draw_p ? 2 : 0) | (draw_s ? 1 : 0));
            A.trace_prof[e] = (uint8_t)(infection_profile(M, u[2 * c] ^ u[(2 * c + 1) % 16]) |
                                        (infection_profile(M, u[2 * c + 1] ^ u[(2 * c + 2) % 16]) << 4));
        }
        if (draw_p) {
            if (A.st_p[v] == kSusceptible) {
                const uint32_t prof = infection_profile(M, u[2 * c] ^ u[(2 * c + 1) % 16]);
                const uint64_t claim = ((uint64_t)(pos * 8u + (uint32_t)c) << 32) | ((uint32_t)f.gen_p << 8) | prof;
                const uint64_t old = atomicMin((unsigned long long*)&A.claim_p[v], (unsigned long long)claim);
                newp[c] = old == ~0ull;
                attempt |= 2;
            }
        }
        if (draw_s) {
            if (A.st_s[v] == kSusceptible) {
                const uint32_t prof = infection_profile(M, u[2 * c + 1] ^ u[(2 * c + 2) % 16]);
                const uint64_t claim = ((uint64_t)(pos * 8u + (uint32_t)c) << 32) | ((uint32_t)f.gen_s << 8) | prof;
                const uint64_t old = atomicMin((unsigned long long*)&A.claim_s[v], (unsigned long long)claim);
                news[c] = old == ~0ull;
                attempt |= 1;
            }
        }
        if (A.trace_attempt) A.trace_attempt[(int64_t)pos * C + c] = attempt;
    }
}

// Appends the victims this lane posted first to the claim lists (as
// id_bias + victim).  All 64 lanes must call.
__device__ __forceinline__ void append_claims(uint32_t* n_claimed, int32_t* __restrict__ list_p,
                                              int32_t* __restrict__ list_s, int32_t id_bias,
                                              const int32_t (&vic)[8], const bool (&newp)[8], const bool (&news)[8],
                                              int lane) {
    uint32_t np = 0, ns = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        np += newp[c];
        ns += news[c];
    }
    uint32_t ip = wave_append(&n_claimed[0], np, lane);
    uint32_t is = wave_append(&n_claimed[1], ns, lane);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        if (newp[c]) list_p[ip++] = id_bias + vic[c];
        if (news[c]) list_s[is++] = id_bias + vic[c];
    }
}

__device__ __forceinline__ void clear_contacts(int32_t (&vic)[8], int (&typ)[8], bool (&newp)[8], bool (&news)[8]) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        vic[c] = -1;
        typ[c] = CT_NONE;
        newp[c] = news[c] = false;
    }
}

// D6w + D7: one infector of kernel_doWeekday (PandemicSim.cu:3236-3308) with
// device_makeContacts_weekday (:1669-1829).  idx: the infector; pos: its
// position in the day's infected list; cbase, eoff: the day's contact and
// errand counters.
__device__ __forceinline__ void weekday_contacts(const Model& M, const AgentArrays& A, int day, uint64_t cbase,
                                                 uint64_t eoff, uint64_t workplace_off, int32_t idx, int64_t pos,
                                                 int32_t (&vic)[8], int (&typ)[8], bool (&newp)[8],
                                                 bool (&news)[8]) {
    const Infectivity f = infectivity(M, A, idx, day);
    if (!can_infect(M, f) && !A.trace_victim) return;
    const uint64_t Mc = cbase + (uint64_t)(6 * pos);
    const U4 ua = draw4(Mc, M.key0, M.key1);
    const U4 ub = draw4(Mc + 1, M.key0, M.key1);
    const uint32_t u[8] = {ua.w[0], ua.w[1], ua.w[2], ua.w[3], ub.w[0], ub.w[1], ub.w[2], ub.w[3]};
    const uint8_t myage = A.age[idx];
    const bool adult = myage == kAgeAdult;
    // household
    const int32_t hh = A.household[idx];
    const int32_t hoff = A.hh_off[hh];
    const int32_t hcnt = A.hh_off[hh + 1] - hoff;
    // workplace (device_recalcWorkplace PandemicSim.cu:3872-3890)
    const U4 uw = draw4(workplace_off + (uint64_t)(idx / 4), M.key0, M.key1);
    uint8_t a2 = myage;
    const int32_t wp = assign_workplace(M, word(uw, idx % 4), &a2);
    const int32_t woff = A.wp_off[wp];
    const int32_t wcnt = A.wp_off[wp + 1] - woff;
    const int wpd = workplace_max_contacts(M, wp);
    // errands (device_recalc_weekdayErrandDests_assignProfile :3667-3705)
    const U4 ue = draw4(eoff + (uint64_t)(idx / 2), M.key0, M.key1);
    const int rs = 2 * (idx % 2);
    int32_t d1, d2;
    weekday_dests(M, word(ue, rs), word(ue, rs + 1), adult, d1, d2);
    const int prof = adult ? (int)((word(ue, rs) ^ word(ue, rs + 1)) % 3u) : 3;
    const int h0 = M.wd_assign[prof][0], h1 = M.wd_assign[prof][1];
    uint32_t k0, k1;
    weekday_keys(M, d1, d2, k0, k1);
    const int32_t e0off = A.er_off[k0], e0cnt = A.er_off[k0 + 1] - e0off;
    const int32_t e1off = A.er_off[k1], e1cnt = A.er_off[k1 + 1] - e1off;
    const int wtype = adult ? CT_WORKPLACE : CT_SCHOOL;
    const int etype = adult ? CT_ERRAND : CT_AFTERSCHOOL;
    const SlotPlan P{hoff, hcnt, woff, wcnt, e0off, e0cnt, e1off, e1cnt, wpd, h0, h1,
                     {u[0], u[1], u[2], u[3], u[4], u[5], u[6], u[7]}};
    const bool resolved = A.trace_victim != nullptr;
    int kval = 0;
    int32_t where[8];  // the contact's location (trace only)
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        int t;
        if (resolved) {
            if (c < 3) {
                t = select_person(idx, hoff, hcnt, u[c], CT_HOME, nullptr, vic[c]);
                where[c] = hh;
            } else if (c < 3 + wpd) {
                t = select_person(idx, woff, wcnt, u[c], wtype, A.wp_people, vic[c]);
                where[c] = wp;
            } else {
                const int e = c - 3 - wpd;
                if (e < h0) {
                    t = select_person(idx, e0off, e0cnt, u[c], etype, A.er_people, vic[c]);
                    where[c] = d1;
                } else if (e < h0 + h1) {
                    t = select_person(idx, e1off, e1cnt, u[c], etype, A.er_people, vic[c]);
                    where[c] = adult ? M.W + d2 : d2;
                } else {
                    vic[c] = -1;
                    t = CT_NONE;
                    where[c] = -1;
                }
            }
        } else {
            // select_person's type without the pick
            const int e = c - 3 - wpd;
            const int32_t cnt = c < 3 ? hcnt : (c < 3 + wpd ? wcnt : (e < h0 ? e0cnt : (e < h0 + h1 ? e1cnt : 0)));
            const int ty = c < 3 ? CT_HOME : (c < 3 + wpd ? wtype : etype);
            t = (e < h0 + h1 && cnt > 1) ? ty : CT_NONE;
        }
        typ[c] = t;
        kval += t != CT_NONE;
    }
    apply_drops(M, day, false, Mc, 8, vic, typ);
    if (A.trace_victim) {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            A.trace_victim[pos * 8 + c] = vic[c];
            A.trace_type[pos * 8 + c] = (uint8_t)typ[c];
            A.trace_loc[pos * 8 + c] = vic[c] < 0 ? -1 : where[c];
        }
    }
    // Traced days run the draws of every infector, as the reference does: an
    // infector that cannot infect has p <= 0 (or NaN), so none of its draws
    // succeeds and it posts no claim.
    if (can_infect(M, f) || A.trace_draw)
        process_contacts<8>(M, A, f, kval, vic, typ, P, resolved, idx, Mc + 2, (uint32_t)pos, newp, news);
    else if (A.trace_attempt)
        for (int c = 0; c < 8; ++c) A.trace_attempt[pos * 8 + c] = 0;
}

// D6e + D7: one infector of kernel_doWeekend (PandemicSim.cu:3311-3379) with
// device_makeContacts_weekend (:1833-1950).
__device__ __forceinline__ void weekend_contacts(const Model& M, const AgentArrays& A, int day, uint64_t cbase,
                                                 uint64_t eoff, int32_t idx, int64_t pos, int32_t (&vic)[8],
                                                 int (&typ)[8], bool (&newp)[8], bool (&news)[8]) {
    const Infectivity f = infectivity(M, A, idx, day);
    if (!can_infect(M, f) && !A.trace_victim) return;
    const uint64_t Mc = cbase + (uint64_t)(6 * pos);
    const U4 u = draw4(Mc, M.key0, M.key1);
    const int32_t hh = A.household[idx];
    const int32_t hoff = A.hh_off[hh];
    const int32_t hcnt = A.hh_off[hh + 1] - hoff;
    const bool resolved = A.trace_victim != nullptr;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (resolved)
            typ[c] = select_person(idx, hoff, hcnt, u.w[c], CT_HOME, nullptr, vic[c]);
        else
            typ[c] = hcnt > 1 ? CT_HOME : CT_NONE;
    }
    const int prof = (int)(u.w[3] % 6u);
    uint32_t v32[2];
    draw2x32((uint32_t)(Mc + 1), (uint32_t)M.key0, (uint32_t)M.key1, v32);
    int32_t hour[3], loc[3];
    weekend_errands(M, eoff + (uint64_t)(int64_t)(2 * idx), hour, loc);
    int32_t where[5] = {hh, hh, hh, -1, -1};  // the contact's location (trace only)
    int32_t eo[2], ec[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int slot = M.we_assign[prof][k];
        const int hsel = slot == 0 ? hour[0] : (slot == 1 ? hour[1] : hour[2]);
        const int lsel = slot == 0 ? loc[0] : (slot == 1 ? loc[1] : loc[2]);
        const int32_t l = min(max(lsel - M.we_base, 0), M.we_E - 1);
        const int32_t key = hsel * M.we_E + l;
        eo[k] = A.er_off[key];
        ec[k] = A.er_off[key + 1] - eo[k];
        if (resolved)
            typ[3 + k] = select_person(idx, eo[k], ec[k], v32[k], CT_ERRAND, A.er_people, vic[3 + k]);
        else
            typ[3 + k] = ec[k] > 1 ? CT_ERRAND : CT_NONE;
        where[3 + k] = hsel * M.W + lsel;
    }
    const SlotPlan P{hoff, hcnt, 0, 0, eo[0], ec[0], eo[1], ec[1], 0, 1, 1,
                     {u.w[0], u.w[1], u.w[2], v32[0], v32[1], 0u, 0u, 0u}};
    int kval = 0;
#pragma unroll
    for (int c = 0; c < 5; ++c) kval += typ[c] != CT_NONE;
    apply_drops(M, day, true, Mc, 5, vic, typ);
    if (A.trace_victim) {
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            A.trace_victim[pos * 5 + c] = vic[c];
            A.trace_type[pos * 5 + c] = (uint8_t)typ[c];
            A.trace_loc[pos * 5 + c] = vic[c] < 0 ? -1 : where[c];
        }
    }
    // Traced days run the draws of every infector, as the reference does: an
    // infector that cannot infect has p <= 0 (or NaN), so none of its draws
    // succeeds and it posts no claim.
    if (can_infect(M, f) || A.trace_draw)
        process_contacts<5>(M, A, f, kval, vic, typ, P, resolved, idx, Mc + 2, (uint32_t)pos, newp, news);
    else if (A.trace_attempt)
        for (int c = 0; c < 5; ++c) A.trace_attempt[pos * 5 + c] = 0;
}

// ---------------- queried-bucket marks ----------------

// Sets bit kbias + key when key < K.
__device__ __forceinline__ void set_mark(uint32_t* __restrict__ mark, int32_t key, int32_t K, uint32_t kbias) {
    if ((uint32_t)key >= (uint32_t)K) return;
    const uint32_t k = kbias + (uint32_t)key;
    const uint32_t bit = 1u << (k & 31);
    if (!(__hip_atomic_load(&mark[k >> 5], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & bit))
        atomicOr(&mark[k >> 5], bit);
}

// The weekday errand buckets an infector draws contacts from: hour h is
// used iff the errand profile asks for contacts in it (the workplace takes at
// most 3 of the 5 non-household slots, so the first requested hour always
// gets one; PandemicSim.cu:1776-1813).
__device__ __forceinline__ void mark_weekday(const Model& M, const AgentArrays& A, int day, uint64_t eoff, int32_t idx,
                                             uint32_t* __restrict__ mark, uint32_t kbias) {
    // The contact kernels also draw for infectors that cannot infect when
    // tracing (their contacts are traced), so their buckets are queried too.
    const Infectivity f = infectivity(M, A, idx, day);
    if (!can_infect(M, f) && !A.trace_victim) return;
    const bool adult = A.age[idx] == kAgeAdult;
    const U4 ue = draw4(eoff + (uint64_t)(idx / 2), M.key0, M.key1);
    const int rs = 2 * (idx % 2);
    int32_t d1, d2;
    weekday_dests(M, word(ue, rs), word(ue, rs + 1), adult, d1, d2);
    uint32_t k0, k1;
    weekday_keys(M, d1, d2, k0, k1);
    const int prof = adult ? (int)((word(ue, rs) ^ word(ue, rs + 1)) % 3u) : 3;
    if (M.wd_assign[prof][0] > 0) set_mark(mark, (int32_t)k0, M.wd_K, kbias);
    if (M.wd_assign[prof][1] > 0) set_mark(mark, (int32_t)k1, M.wd_K, kbias);
}

// The two weekend errand buckets of an infector (PandemicSim.cu:1891-1946).
__device__ __forceinline__ void mark_weekend(const Model& M, const AgentArrays& A, int day, uint64_t cbase,
                                             uint64_t eoff, int64_t pos, int32_t idx, uint32_t* __restrict__ mark,
                                             uint32_t kbias) {
    const Infectivity f = infectivity(M, A, idx, day);
    if (!can_infect(M, f) && !A.trace_victim) return;  // as mark_weekday
    const U4 u = draw4(cbase + (uint64_t)(6 * pos), M.key0, M.key1);
    const int prof = (int)(u.w[3] % 6u);
    int32_t hour[3], loc[3];
    weekend_errands(M, eoff + (uint64_t)(int64_t)(2 * idx), hour, loc);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int slot = M.we_assign[prof][k];
        const int hsel = slot == 0 ? hour[0] : (slot == 1 ? hour[1] : hour[2]);
        const int lsel = slot == 0 ? loc[0] : (slot == 1 ? loc[1] : loc[2]);
        set_mark(mark, hsel * M.we_E + (lsel - M.we_base), kWeHours * M.we_E, kbias);
    }
}

// ---------------- kernels used by both engines ----------------
// (internal linkage: each translation unit compiles its own copy)
namespace {

// Compaction of per-block staged runs (counts cnt, offsets off) into one list.
__global__ void k_compact_runs(const int32_t* __restrict__ skip, const uint32_t* __restrict__ stage_keys,
                               const int32_t* __restrict__ stage_ids, const uint32_t* __restrict__ blk_cnt,
                               const uint32_t* __restrict__ blk_off, uint32_t* __restrict__ sel_keys,
                               int32_t* __restrict__ sel_ids, int stage_stride, uint32_t* __restrict__ n_sel,
                               int nblk) {
    if (*skip) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) *n_sel = blk_off[nblk - 1] + blk_cnt[nblk - 1];
    const uint32_t n = blk_cnt[blockIdx.x];
    const uint32_t off = blk_off[blockIdx.x];
    const int64_t src = (int64_t)blockIdx.x * stage_stride;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        sel_keys[off + i] = stage_keys[src + i];
        sel_ids[off + i] = stage_ids[src + i];
    }
}

// Winners write status = profile, day = current_day + 1, generation
// (device_setInfectionStatus PandemicSim.cu:3048-3067); claim words reset.
__global__ void k_apply(const DayState* __restrict__ st, uint64_t* __restrict__ claim_p,
                        uint64_t* __restrict__ claim_s, const int32_t* __restrict__ list_p,
                        const int32_t* __restrict__ list_s, uint8_t* __restrict__ st_p, uint8_t* __restrict__ st_s,
                        uint8_t* __restrict__ dy_p, uint8_t* __restrict__ dy_s, uint8_t* __restrict__ gn_p,
                        uint8_t* __restrict__ gn_s) {
    if (st->done) return;
    const uint32_t np = st->n_claimed[0], ns = st->n_claimed[1];
    const uint8_t dset = (uint8_t)(st->day + 1);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (int64_t)np + ns; i += stride) {
        if (i < np) {
            const int32_t v = list_p[i];
            const uint64_t c = claim_p[v];
            st_p[v] = (uint8_t)(c & 0xFF);
            dy_p[v] = dset;
            gn_p[v] = (uint8_t)((c >> 8) & 0xFF);
            claim_p[v] = ~0ull;
        } else {
            const int32_t v = list_s[i - np];
            const uint64_t c = claim_s[v];
            st_s[v] = (uint8_t)(c & 0xFF);
            dy_s[v] = dset;
            gn_s[v] = (uint8_t)((c >> 8) & 0xFF);
            claim_s[v] = ~0ull;
        }
    }
}

// Plain copy of n16 16-byte words (snapshot / restore of the agent state).
__global__ void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

// setup_initialInfected (PandemicSim.cu:708-758).
__global__ void k_seed_infections(const int32_t* __restrict__ ids, int n_p, int n_s, uint8_t* st_p, uint8_t* st_s,
                                  uint8_t* dy_p, uint8_t* dy_s, uint8_t* gn_p, uint8_t* gn_s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_p + n_s) return;
    const int id = ids[i];
    if (i < n_p) {
        st_p[id] = kInfected;
        dy_p[id] = 0;
        gn_p[id] = 0;
    } else {
        st_s[id] = kInfected;
        dy_s[id] = 0;
        gn_s[id] = 0;
    }
}

// Generation histograms (final_countReproduction, PandemicSim.cu:2236-2260).
__global__ void k_gen_hist(const uint8_t* __restrict__ gn_p, const uint8_t* __restrict__ gn_s, int64_t N,
                           unsigned long long* __restrict__ hist) {
    __shared__ uint32_t s_h[512];
    for (int i = threadIdx.x; i < 512; i += blockDim.x) s_h[i] = 0;
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        atomicAdd(&s_h[gn_p[i]], 1u);
        atomicAdd(&s_h[256 + gn_s[i]], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 512; i += blockDim.x)
        if (s_h[i]) atomicAdd(&hist[i], (unsigned long long)s_h[i]);
}

}  // namespace
}  // namespace pt
