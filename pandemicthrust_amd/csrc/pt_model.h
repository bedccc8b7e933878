// pt_model.h — model tables and per-agent device functions of the daily loop.
//
// The functions here are the per-thread bodies of the reference's device
// functions (citations relative to PandemicThrust/PandemicThrust/), written
// for gfx950.  Floating point follows the reference's single-precision
// operation order; the library is built with -ffp-contract=off and IEEE
// division so that results are bit-identical to the reference semantics.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pt {

constexpr int kTypes = 14;
constexpr int kHhRows = 9;
constexpr int kChildRows = 5;
constexpr int kAgeAdult = 5;
constexpr int kCulmination = 10;
constexpr int kFirstErrandRow = 9;
constexpr int kTypeAfterschool = 8;
constexpr int kTypePreschool = 3;
constexpr int kWdHours = 2, kWeHours = 10;
constexpr int kWdErrands = 2, kWeErrands = 3;
constexpr int kContactsWd = 8, kContactsWe = 5;
enum ContactType : int { CT_NONE = 0, CT_WORKPLACE, CT_SCHOOL, CT_ERRAND, CT_AFTERSCHOOL, CT_HOME };
constexpr uint8_t kRecovered = 0, kSusceptible = 1, kInfected = 2;
constexpr uint8_t kDayNone = 255, kGenNone = 255;

// Model tables resident in device memory, one copy per simulation handle.
struct Model {
    float hh_cdf[kHhRows];
    int32_t hh_adults[kHhRows];
    int32_t hh_children[kHhRows];
    float child_cdf[kChildRows];
    float wp_pdf[kTypes];
    int32_t wp_count[kTypes];
    int32_t wp_offset[kTypes];
    int32_t wp_max[kTypes];
    float wd_pdf[kTypes];
    float we_pdf[kTypes];
    int32_t wd_assign[4][2];
    int32_t we_assign[6][2];
    float shed[6][10];
    float factor[2];
    float pct;
    int32_t pad0;
    uint64_t key0, key1;  // Threefry2x64 key {SEED[0], SEED[1]}
    int32_t N, H, W;
    // Dense errand-table layout: every errand destination lies in
    // [base, W), so bucket (hour, loc) maps to hour * E + (loc - base).
    // Weekday hour 1 keeps only [wd_base1, W) (weekday_keys).
    int32_t wd_base, wd_E;
    int32_t we_base, we_E;
    int32_t wd_base1, wd_E1, wd_K;
    int32_t pad1;
    // Interventions (none in the reference; the neutral values reproduce it
    // exactly): per strain, infectious while lat <= day - infection day <
    // lat + dur (reference: 0, 10); weekday school closure [school_start,
    // school_start + school_days); shares of workplace / errand contacts
    // dropped; share of agents immune to each strain from setup.  Drop and
    // vaccination draws come from their own Threefry stream (key ikey0,
    // ikey1), so no draw of the model's stream is consumed or moved.
    int32_t lat[2], dur[2];
    int32_t school_start, school_days;
    float wp_red, er_red, vacc[2];
    int32_t drops;  // any contact can be dropped (school closure or a reduction > 0)
    int32_t pad2;
    uint64_t ikey0, ikey1;
};

// ---------------- Threefry (Salmon et al., SC'11) ----------------
__host__ __device__ __forceinline__ uint64_t rotl64(uint64_t x, unsigned n) {
    return (x << n) | (x >> (64 - n));
}
__host__ __device__ __forceinline__ uint32_t rotl32(uint32_t x, unsigned n) {
    return (x << n) | (x >> (32 - n));
}

// Threefry2x64, 20 rounds, counter {c, c}: the reference always passes the
// same 64-bit offset in both counter words (e.g. PandemicSim.cu:2340).
struct U4 { uint32_t w[4]; };

// Word i (0..3) of a draw, selected without dynamic indexing (an indexed
// read of a register array is placed in scratch memory).
__host__ __device__ __forceinline__ uint32_t word(const U4& u, int i) {
    return i == 0 ? u.w[0] : (i == 1 ? u.w[1] : (i == 2 ? u.w[2] : u.w[3]));
}
// 64-bit rotate left by a constant 0 < n < 64 as two 32-bit funnel shifts
// (v_alignbit_b32): hi' = (hi:lo) >> (32 - n), lo' = (lo:hi) >> (32 - n)
// for n < 32; halves swapped first for n >= 32.
template <unsigned n>
__device__ __forceinline__ uint64_t rotl64c(uint64_t x) {
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    if (n == 32) return ((uint64_t)lo << 32) | hi;
    const uint32_t a = n < 32 ? hi : lo, b = n < 32 ? lo : hi;
    constexpr unsigned s = 32u - (n & 31u);
    return ((uint64_t)__builtin_amdgcn_alignbit(a, b, s) << 32) | __builtin_amdgcn_alignbit(b, a, s);
}

__device__ __forceinline__ U4 draw4(uint64_t c, uint64_t k0, uint64_t k1) {
    const uint64_t k2 = 0x1BD11BDAA9FC1A22ull ^ k0 ^ k1;
    uint64_t x0 = c + k0, x1 = c + k1;
#define PT_R(n) x0 += x1; x1 = rotl64c<n>(x1); x1 ^= x0;
#define PT_INJ(a, b, s) x0 += a; x1 += b + (uint64_t)(s);
    PT_R(16) PT_R(42) PT_R(12) PT_R(31) PT_INJ(k1, k2, 1)
    PT_R(16) PT_R(32) PT_R(24) PT_R(21) PT_INJ(k2, k0, 2)
    PT_R(16) PT_R(42) PT_R(12) PT_R(31) PT_INJ(k0, k1, 3)
    PT_R(16) PT_R(32) PT_R(24) PT_R(21) PT_INJ(k1, k2, 4)
    PT_R(16) PT_R(42) PT_R(12) PT_R(31) PT_INJ(k2, k0, 5)
#undef PT_R
#undef PT_INJ
    U4 r;
    // Union layout of the reference (u32 word 2i = low half of output i).
    r.w[0] = (uint32_t)x0;
    r.w[1] = (uint32_t)(x0 >> 32);
    r.w[2] = (uint32_t)x1;
    r.w[3] = (uint32_t)(x1 >> 32);
    return r;
}

// Threefry2x32, 20 rounds (the weekend errand-contact draw,
// PandemicSim.cu:1895-1901).
__device__ __forceinline__ void draw2x32(uint32_t c, uint32_t k0, uint32_t k1, uint32_t out[2]) {
    const uint32_t k2 = 0x1BD11BDAu ^ k0 ^ k1;
    uint32_t x0 = c + k0, x1 = c + k1;
#define PT_R(n) x0 += x1; x1 = rotl32(x1, n); x1 ^= x0;
#define PT_INJ(a, b, s) x0 += a; x1 += b + (uint32_t)(s);
    PT_R(13) PT_R(15) PT_R(26) PT_R(6) PT_INJ(k1, k2, 1)
    PT_R(17) PT_R(29) PT_R(16) PT_R(24) PT_INJ(k2, k0, 2)
    PT_R(13) PT_R(15) PT_R(26) PT_R(6) PT_INJ(k0, k1, 3)
    PT_R(17) PT_R(29) PT_R(16) PT_R(24) PT_INJ(k1, k2, 4)
    PT_R(13) PT_R(15) PT_R(26) PT_R(6) PT_INJ(k2, k0, 5)
#undef PT_R
#undef PT_INJ
    out[0] = x0;
    out[1] = x1;
}

// (float)u / UNSIGNED_MAX: the divisor converts to 2^32 in float, so this
// is an exact scaling of the round-to-nearest conversion.
__device__ __forceinline__ float unit_float(uint32_t u) { return (float)u / 4294967296.0f; }

// ---------------- setup draws ----------------
// device_setup_fishHouseholdType (PandemicSim.cu:2360-2371)
__device__ __forceinline__ int fish_household(const Model& M, uint32_t r) {
    const float y = unit_float(r);
    int row = 0;
    while (y > M.hh_cdf[row] && row < kHhRows - 1) row++;
    return row;
}

// device_setup_fishWorkplace (PandemicSim.cu:2374-2399)
__device__ __forceinline__ int32_t fish_workplace(const Model& M, uint32_t r) {
    float y = unit_float(r);
    int row = 0;
    while (M.wp_pdf[row] < y && row < kTypes - 1) {
        y -= M.wp_pdf[row];
        row++;
    }
    const float frac = y / M.wp_pdf[row];
    const int type_count = M.wp_count[row];
    int bn = (int)(frac * (float)type_count);
    if (bn >= type_count) bn = type_count - 1;
    return bn + M.wp_offset[row];
}

// device_setup_fishSchoolAndAge (PandemicSim.cu:2402-2442)
__device__ __forceinline__ int32_t fish_school_and_age(const Model& M, uint32_t r, uint8_t* age) {
    const float y = unit_float(r);
    int row = 0;
    while (row < kChildRows - 1 && y > M.child_cdf[row]) row++;
    const int wp_type = row + kTypePreschool;
    float frac;
    if (row == 0) {
        frac = y / M.child_cdf[0];
    } else {
        const float pdf_here = M.child_cdf[row] - M.child_cdf[row - 1];
        const float y_here = y - M.child_cdf[row - 1];
        frac = y_here / pdf_here;
    }
    const int type_count = M.wp_count[wp_type];
    int bn = (int)(frac * (float)type_count);
    if (bn >= type_count) bn = type_count - 1;
    *age = (uint8_t)row;
    return bn + M.wp_offset[wp_type];
}

// device_setup_assignWorkplaceOrSchool (PandemicSim.cu:2445-2457)
__device__ __forceinline__ int32_t assign_workplace(const Model& M, uint32_t r, uint8_t* age) {
    if (*age == kAgeAdult) return fish_workplace(M, r);
    return fish_school_and_age(M, r, age);
}

// ---------------- errand draws ----------------
// The errand walks below run over rows kFirstErrandRow..kTypes-1 with every
// step unrolled and predicated: each step subtracts the row's pdf only where
// the reference's loop would continue, so the float operations and their
// order are exactly the loop's, while the table reads are uniform (constant
// row indices) instead of per-lane gathers inside a divergent loop.
static_assert(kTypes - kFirstErrandRow == 5, "errand walks are unrolled over 5 rows");

struct RowPick {
    float frac;   // the walk's remaining uniform
    float pdf;    // pdf of the row it stopped at
    int count;    // locations of that row
    int offset;   // first location of that row
};

// while (frac > pdf[row] && row < kTypes - 1) { frac -= pdf[row]; row++; }
// from row kFirstErrandRow, over the pdf row table `pdf`.
__device__ __forceinline__ RowPick walk_errand_rows(const Model& M, const float* pdf, float frac) {
    RowPick r{frac, pdf[kFirstErrandRow], M.wp_count[kFirstErrandRow], M.wp_offset[kFirstErrandRow]};
    bool go = true;
#pragma unroll
    for (int row = kFirstErrandRow; row < kTypes - 1; ++row) {
        go = go && r.frac > pdf[row];
        const float rest = r.frac - pdf[row];
        r.frac = go ? rest : r.frac;
        r.pdf = go ? pdf[row + 1] : r.pdf;
        r.count = go ? M.wp_count[row + 1] : r.count;
        r.offset = go ? M.wp_offset[row + 1] : r.offset;
    }
    return r;
}

// device_fishAfterschoolOrErrandDestination_weekday (PandemicSim.cu:2909-2947)
// Every lane walks the rows and divides, and the child's after-school pick
// is selected afterwards: a wave holds adults and children alike, so a
// branch would run both sides anyway, plus its mask handling.
__device__ __forceinline__ int32_t fish_weekday_dest(const Model& M, uint32_t r, bool adult) {
    const float u = unit_float(r);
    const RowPick p = walk_errand_rows(M, M.wd_pdf, u);
    const float q = p.frac / p.pdf;
    const float frac = adult ? q : u;
    const int type_count = adult ? p.count : M.wp_count[kTypeAfterschool];
    const int type_offset = adult ? p.offset : M.wp_offset[kTypeAfterschool];
    uint32_t bn = (uint32_t)(frac * (float)type_count);
    if (bn >= (uint32_t)type_count) bn = (uint32_t)(type_count - 1);
    bn += (uint32_t)type_offset;
    return (int32_t)bn;
}

// device_fishWeekendErrandDestination (PandemicSim.cu:1017-1044)
__device__ __forceinline__ int32_t fish_weekend_dest(const Model& M, uint32_t r) {
    const RowPick p = walk_errand_rows(M, M.we_pdf, unit_float(r));
    const float frac = p.frac / p.pdf;
    int bn = (int)(frac * (float)p.count);
    if (bn >= p.count) bn = p.count - 1;
    return bn + p.offset;
}

// Weekday errand destinations of one person from its two draws
// (device_assignAfterschoolOrErrandDests_weekday, PandemicSim.cu:2950-2969),
// returned as locations (hour 1's W offset is not added).
__device__ __forceinline__ void weekday_dests(const Model& M, uint32_t r1, uint32_t r2, bool adult, int32_t& d1,
                                              int32_t& d2) {
    d1 = fish_weekday_dest(M, r1, adult);
    const int32_t e2 = fish_weekday_dest(M, r2, true);
    d2 = adult ? e2 : d1;
}

// ---------------- interventions ----------------

// Sixteen-bit uniform of contact slot c (0..7) from one draw of the
// intervention stream.
__host__ __device__ __forceinline__ float drop_uniform(const U4& x, int c) {
    return (float)((x.w[c >> 1] >> (16 * (c & 1))) & 0xFFFFu) * (1.0f / 65536.0f);
}

// Contact interventions for one infector, applied after its contacts are
// chosen: on a weekday inside the closure window school and after-school
// contacts are dropped; a workplace (errand) contact is dropped when its
// slot's uniform is below the workplace (errand) reduction.  A dropped slot
// becomes (victim -1, CT_NONE); the infector's normalising contact count is
// left as it was, so closing a place removes its transmission rather than
// moving it to the remaining contacts.  Mc is the infector's contact counter.
__device__ __forceinline__ void apply_drops(const Model& M, int day, bool weekend, uint64_t Mc, int C,
                                            int32_t (&vic)[8], int (&typ)[8]) {
    if (!M.drops) return;
    const U4 x = draw4(Mc, M.ikey0, M.ikey1);
    const bool closed = !weekend && day >= M.school_start && day < M.school_start + M.school_days;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        if (c >= C) break;
        const int t = typ[c];
        bool drop = false;
        if (t == CT_SCHOOL || t == CT_AFTERSCHOOL)
            drop = closed;
        else if (t == CT_WORKPLACE)
            drop = drop_uniform(x, c) < M.wp_red;
        else if (t == CT_ERRAND)
            drop = drop_uniform(x, c) < M.er_red;
        if (drop) {
            vic[c] = -1;
            typ[c] = CT_NONE;
        }
    }
}

// Counter of agent i's vaccination draw in the intervention stream (kept
// apart from the contact counters by the top bit).
__host__ __device__ __forceinline__ uint64_t vaccination_counter(int64_t i) { return (1ull << 63) | (uint64_t)i; }

// Weekday errand-table keys of one person.  Hour 0 spans every destination
// in [wd_base, W): key d1 - wd_base.  In hour 1 a child stays at its
// after-school site and makes no contacts there (contact profile 3 has none
// in hour 1, PandemicSim.cu:454-455), and adults only reach [wd_base1, W)
// (wd_base1 = the lowest destination fish_weekday_dest can give an adult), so
// no infector ever draws from an hour-1 bucket below wd_base1.  Hour 1 keeps
// [wd_base1, W) at keys wd_E + (d2 - wd_base1); every hour-1 entry below it
// goes to the last bucket wd_K - 1, which nobody queries.  Returns false for
// a destination outside [wd_base, W) (the keys are then clamped in range).
// (Bitwise & and selects, not && and ?: over loads: the compiler would
// otherwise branch around the model reads.)
__device__ __forceinline__ bool weekday_keys(const Model& M, int32_t d1, int32_t d2, uint32_t& k0, uint32_t& k1) {
    const int32_t E = M.wd_E, E1 = M.wd_E1, K = M.wd_K, base = M.wd_base, base1 = M.wd_base1, W = M.W;
    const int32_t l1 = d1 - base, l2 = d2 - base1;
    const bool ok = ((uint32_t)l1 < (uint32_t)E) & (d2 >= base) & (d2 < W);
    k0 = (uint32_t)min(max(l1, 0), E - 1);
    const uint32_t hour1 = (uint32_t)(E + min(l2, E1 - 1));
    k1 = l2 >= 0 ? hour1 : (uint32_t)(K - 1);
    return ok;
}

// device_generateWeekendErrands (PandemicSim.cu:980-1014): three distinct
// hours and their destinations.
__device__ __forceinline__ void weekend_errands(const Model& M, uint64_t off, int32_t hour[3], int32_t loc[3]) {
    const U4 a = draw4(off, M.key0, M.key1);
    const U4 b = draw4(off + 1, M.key0, M.key1);
    const int h1 = (int)(a.w[0] % kWeHours);
    int h2 = (int)(a.w[1] % kWeHours);
    if (h2 == h1) h2 = (h2 + 1) % kWeHours;
    int h3 = (int)(a.w[2] % kWeHours);
    // while (h3 == h1 || h3 == h2) h3 = (h3 + 1) % kWeHours, unrolled: two
    // forbidden hours end it within two steps.
    h3 = (h3 == h1 || h3 == h2) ? (h3 + 1) % kWeHours : h3;
    h3 = (h3 == h1 || h3 == h2) ? (h3 + 1) % kWeHours : h3;
    hour[0] = h1;
    hour[1] = h2;
    hour[2] = h3;
    loc[0] = fish_weekend_dest(M, a.w[3]);
    loc[1] = fish_weekend_dest(M, b.w[0]);
    loc[2] = fish_weekend_dest(M, b.w[1]);
}

// device_getWorkplaceMaxContacts (PandemicSim.cu:3492-3511)
__device__ __forceinline__ int workplace_max_contacts(const Model& M, int32_t loc) {
    const int id = loc % M.W;
    int t = 0, off, cnt, mx;
    do {
        off = M.wp_offset[t];
        cnt = M.wp_count[t];
        mx = M.wp_max[t];
        t++;
    } while (id >= off + cnt && t < kTypes);
    return mx;
}

// device_calculateInfectionProbability (PandemicSim.cu:2307-2326)
__device__ __forceinline__ float infection_probability(const Model& M, int profile, int dd, int strain, int kval_sum) {
    if (kval_sum == 0) kval_sum = 1;
    const float alpha = M.shed[profile][dd];
    const float avg = alpha * M.factor[strain];
    return avg / (float)kval_sum;
}

// device_getInfectionProfile (PandemicSim.cu:3974-3988)
__device__ __forceinline__ uint32_t infection_profile(const Model& M, uint32_t r) {
    uint32_t prof = r % 3u + kInfected;
    if (unit_float(r) > M.pct) prof += 3;
    return prof;
}

}  // namespace pt
