// pt_table_noalloc.hip — the allocation-free location-table build
// (C ABI of pt_location_table_noalloc, include/pandemic/pt_abi.h).
//
// Same build and results as pt_location_table (pt_table.hip: the reference's
// cub::DeviceRadixSort::SortPairs + thrust::lower_bound,
// PandemicSim.cu:1212-1248), but the radix scratch lives in a file-scope
// device array and the call allocates nothing, for callers that cannot
// allocate (code under HIP graph capture, tight serving loops).  The array
// holds kNoAllocMaxPairs pairs; every call on one device shares that
// device's copy, so calls that overlap on one device (different streams, or
// different host threads) race and must be serialised by the caller.  Its
// contents are left from the previous call: every scratch word the build
// reads is written earlier in the same build (digit histograms by the
// upsweeps, scan partials by the scans, pairs by the previous pass), so
// nothing needs clearing.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/pandemic/pt_abi.h"
#include "pt_day.h"
#include "pt_sort.h"

using namespace pt;

namespace {

constexpr int64_t kNoAllocMaxPairs = PT_LOCATION_TABLE_NOALLOC_MAX_PAIRS;
constexpr int kNoAllocMaxBits = 8;  // widest digit per radix pass
// digit histograms of at most 1024 radix blocks of 2^8 digits, plus the total
constexpr int64_t kHistWords = ((int64_t)1 << kNoAllocMaxBits) * 1024 + 1;
// scan partial sums: one per 8192 scanned values, plus slack
constexpr int64_t kScanWords = (kHistWords > kNoAllocMaxPairs ? kHistWords : kNoAllocMaxPairs) / 8192 + 4;

constexpr int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }
constexpr int64_t kOffKeys = 0;
constexpr int64_t kOffPairsA = align256(kOffKeys + kNoAllocMaxPairs * 4);
constexpr int64_t kOffPairsB = align256(kOffPairsA + kNoAllocMaxPairs * 8);
constexpr int64_t kOffHist = align256(kOffPairsB + kNoAllocMaxPairs * 8);
constexpr int64_t kOffScan = align256(kOffHist + kHistWords * 4);
constexpr int64_t kOffKept = align256(kOffScan + kScanWords * 4);
constexpr int64_t kScratchBytes = align256(kOffKept + 4);

// One copy per device, kept until the process ends.  Host code never passes
// this name to a kernel or a copy: it uses the device address that
// hipGetSymbolAddress returns (scratch_ptr).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (current), looked up on the
// first call on that device.
int scratch_ptr(int device, unsigned char** out) {
    if (device < 0 || device >= kMaxDevices) return fail(PT_EINVAL, "device %d outside [0, %d)", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(PT_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        g_scratch_ptr[device] = (unsigned char*)p;
    }
    *out = g_scratch_ptr[device];
    return PT_OK;
}

}  // namespace

extern "C" int pt_location_table_noalloc(const uint32_t* keys, int64_t n, int32_t n_buckets, int32_t* people,
                                         int32_t* offsets, int32_t device, void* stream) {
    if (int rc = location_table_args(keys, n, n_buckets, people, offsets)) return rc;
    if (n > kNoAllocMaxPairs)
        return fail(PT_EINVAL, "%lld pairs exceed the allocation-free limit of %lld", (long long)n,
                    (long long)kNoAllocMaxPairs);
    DeviceScope ds(device);
    HIP_TRY(ds.err);
    unsigned char* base = nullptr;
    if (int rc = scratch_ptr(device, &base)) return rc;
    TableScratch ts;
    ts.keys_b = (uint32_t*)(base + kOffKeys);
    ts.pairs_a = (uint2*)(base + kOffPairsA);
    ts.pairs_b = (uint2*)(base + kOffPairsB);
    ts.hist = (uint32_t*)(base + kOffHist);
    ts.scan_part = (uint32_t*)(base + kOffScan);
    ts.n_kept = (uint32_t*)(base + kOffKept);
    ts.n_cap = kNoAllocMaxPairs;
    ts.max_bits = kNoAllocMaxBits;
    const hipError_t e = build_table(ts, const_cast<uint32_t*>(keys), nullptr, n, nullptr, n_buckets, ID_IDENTITY,
                                     (int32_t)n, people, offsets, nullptr, (hipStream_t)stream, nullptr);
    if (e != hipSuccess) return fail(PT_EDEVICE, "location table: %s", hipGetErrorString(e));
    return PT_OK;
}
