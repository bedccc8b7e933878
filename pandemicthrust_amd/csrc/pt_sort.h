// pt_sort.h — stable location tables on gfx950.
//
// Replaces the reference's cub::DeviceRadixSort::SortPairs +
// thrust::lower_bound pairs (PandemicSim.cu:1212-1248, :1316-1354,
// :3835-3862).  Input: n uint32 bucket keys < K whose person ids are implicit
// in the input position (ids ascend within each key in input order).
// Output: people[n] (ids grouped by key, ascending within a key) and
// offsets[K+1] with offsets[k] = number of keys < k.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pt {

// Kernel classes timed with HIP events (pt_sim_kernel_times).
enum KernelClass : int {
    KC_SWEEP = 0,
    KC_PLAN,
    KC_COMPACT,
    KC_GEN_WEEKDAY,
    KC_GEN_WEEKEND,
    KC_SORT_UPSWEEP,
    KC_SORT_SCAN,
    KC_SORT_DOWNSWEEP,
    KC_OFFSETS,
    KC_CONTACTS_WEEKDAY,
    KC_CONTACTS_WEEKEND,
    KC_APPLY,
    KC_END_DAY,
    KC_MARK,
    KC_SELECT,
    KC_SUBSET_SORT,  // idle: subset and full errand tables share one build (the sort classes); index kept
    KC_SETUP,
    KC_COUNT
};

// Optional launch timer.  Timed launches carry their start/stop events in
// the dispatch packet (hipExtLaunchKernel), so timing adds no barrier
// packets between kernels.
struct LaunchTimer {
    virtual bool want(int cls) const = 0;
    // A fresh (start, stop) event pair accounted to class cls.
    virtual void events(int cls, hipEvent_t* start, hipEvent_t* stop) = 0;
    virtual ~LaunchTimer() {}
};

template <typename F, typename... Args>
inline void launch(LaunchTimer* tm, int cls, F kernel, dim3 grid, dim3 block, hipStream_t stream, Args... args) {
    if (tm && tm->want(cls)) {
        hipEvent_t a, b;
        tm->events(cls, &a, &b);
        hipExtLaunchKernelGGL(kernel, grid, block, 0, stream, a, b, 0, args...);
    } else {
        kernel<<<grid, block, 0, stream>>>(args...);
    }
}

// How a person id follows from the input position j.
enum IdMode : int {
    ID_IDENTITY = 0,    // id = j                  (workplace table)
    ID_HOUR_MAJOR = 1,  // id = j mod N            (weekday: hour 0 block, then hour 1 block)
    ID_DIV3 = 2,        // id = j / 3              (weekend: 3 errands per person)
};

struct TableScratch {
    uint32_t* keys_b = nullptr;  // n_cap: the sorted keys of the last pass
    uint2* pairs_a = nullptr;    // n_cap (key, id) pairs between passes
    uint2* pairs_b = nullptr;    // n_cap
    uint32_t* hist = nullptr;    // 256 * kMaxSortBlocks
    uint32_t* scan_part = nullptr;
    uint32_t* n_kept = nullptr;  // pairs kept by a first pass that dropped a key (device)
    int64_t n_cap = 0;
    // Keys come in long runs of equal high bits (tables of many simulations
    // side by side): count digits with wave-aggregated atomics.
    bool grouped_keys = false;
    // Radix passes of at least this many bits (at most 8) run 1024-thread
    // blocks on 8192-pair tiles, narrower ones 512-thread blocks on
    // 4096-pair tiles; 9 keeps every pass at 512 threads.
    int wide_bits = 8;
    // Widest digit of a radix pass (4..11): a table of K buckets takes
    // ceil(log2 K / max_bits) passes.
    int max_bits = 8;
};

int table_scratch_alloc(TableScratch& s, int64_t n_cap);
void table_scratch_free(TableScratch& s);

constexpr uint32_t kNoDrop = 0xFFFFFFFFu;

// A second input for a table whose input is chosen on the device: while
// *flag is nonzero the first radix pass reads keys with the explicit ids vals
// (*n_dev pairs, none dropped) instead of the build's own keys; the later
// passes and the offsets are the same launches either way, so a caller that
// decides between two tables on the device enqueues one build instead of
// two, each idle on the days the other runs.
struct AltInput {
    const int32_t* flag = nullptr;
    const uint32_t* keys = nullptr;
    const int32_t* vals = nullptr;
    const uint32_t* n_dev = nullptr;
};

// Builds the table.  keys is only read.  vals (may be null) holds explicit
// ids in input order; when null the ids follow from mode.  When n_dev is
// set it holds the pair count on the device and n is only the capacity.
// done (device int, may be null) turns every kernel into a no-op when set.
// drop_key = K - 1 leaves the pairs of the last bucket out of the table (a
// bucket nobody queries): that bucket comes out empty, offsets[K - 1] =
// offsets[K] = pairs kept, and ts.n_kept holds that count on the device.
// alt (may be null; requires vals and n_dev null) is the input the device
// may pick instead (AltInput); ts.n_kept then holds the pair count of
// whichever input was read, and n must also bound *alt->n_dev.
hipError_t build_table(TableScratch& s, uint32_t* keys, const int32_t* vals, int64_t n, const uint32_t* n_dev,
                       int32_t K, IdMode mode, int32_t N, int32_t* people_out, int32_t* offsets_out,
                       const int32_t* done, hipStream_t stream, LaunchTimer* tm, uint32_t drop_key = kNoDrop,
                       const AltInput* alt = nullptr);

// Segmented tables: the tables of R independent simulations built by one
// sort whose passes cover only the bits of the largest local key.  Segment r
// holds its pairs at [seg_start[r], seg_start[r] + seg_count[r]) with keys
// local to it (< seg_K[r]) and is sorted in place; seg_start values are
// multiples of 4 and the segments lie in order.  Radix block b works inside
// segment blk_seg[b] only; segment r owns the blocks [seg_blk0[r],
// seg_blk0[r] + seg_nblk[r]) of the kSegBlocks blocks of a pass.  The
// offsets are global: offsets[seg_kbase[r] + q] for q in [0, seg_K[r]] are
// positions of the shared people array; buckets between seg_kbase[r] +
// seg_K[r] and seg_kbase[r + 1] hold no pairs.  Requires seg_kbase[r] +
// seg_K[r] < seg_kbase[r + 1]: the last of those buckets spans the segment's
// alignment gap (unwritten positions) and must not be queried.
constexpr int kSegBlocks = 1024;
struct SegLayout {
    const int32_t* blk_seg = nullptr;     // kSegBlocks
    const int32_t* seg_blk0 = nullptr;    // R
    const int32_t* seg_nblk = nullptr;    // R
    const uint32_t* seg_start = nullptr;  // R (device, per build)
    const uint32_t* seg_count = nullptr;  // R (device, per build)
    const uint32_t* seg_kbase = nullptr;  // R + 1, seg_kbase[R] = the total bucket count
    const int32_t* seg_K = nullptr;       // R
    int R = 0;
};

// Builds every segment's table (ids from vals, explicit).  K_local: the
// largest seg_K; K_total = seg_kbase[R].  n_cap bounds the positions used.
hipError_t build_table_segmented(TableScratch& s, uint32_t* keys, const int32_t* vals, int64_t n_cap,
                                 int32_t K_local, int32_t K_total, const SegLayout& seg, int32_t* people_out,
                                 int32_t* offsets_out, const int32_t* done, hipStream_t stream, LaunchTimer* tm);

// Number of radix passes build_table uses for K buckets at max_bits per pass.
int radix_passes(int32_t K, int max_bits);

// Exclusive scan of n uint32 values (in and out may alias); part needs
// ceil(n / 8192) + 1 entries.
hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* part, const int32_t* done,
                              hipStream_t stream, LaunchTimer* tm = nullptr, int cls = 0);

}  // namespace pt
