// pt_table.hip — stand-alone location-table build on gfx950 (C ABI of
// pt_location_table, include/pandemic/pt_abi.h).
//
// A location table is what the reference builds each day with
// cub::DeviceRadixSort::SortPairs followed by thrust::lower_bound
// (PandemicSim.cu:1212-1248, :1316-1354, :3835-3862): the ids of the pairs
// grouped by bucket key, ascending within a key (the sort is stable and the
// ids come in ascending order), and offsets[k] = the number of keys below k.
// The build is the engine's own radix table build (pt_sort.hip); its radix
// scratch is allocated for each call and freed before the call returns.
#include <hip/hip_runtime.h>

#include "../../include/pandemic/pt_abi.h"
#include "pt_day.h"
#include "pt_sort.h"

using namespace pt;

namespace pt {

int location_table_args(const uint32_t* keys, int64_t n, int32_t n_buckets, const int32_t* people,
                        const int32_t* offsets) {
    if (n < 0 || n >= ((int64_t)1 << 31)) return fail(PT_EINVAL, "pair count %lld outside [0, 2^31)", (long long)n);
    if (n_buckets < 1 || n_buckets == 0x7FFFFFFF)
        return fail(PT_EINVAL, "bucket count %d outside [1, 2^31 - 1)", n_buckets);
    if (!offsets || (n > 0 && (!keys || !people))) return fail(PT_EINVAL, "null device pointer");
    return PT_OK;
}

}  // namespace pt

extern "C" int pt_location_table(const uint32_t* keys, int64_t n, int32_t n_buckets, int32_t* people,
                                 int32_t* offsets, int32_t device, void* stream) {
    if (int rc = location_table_args(keys, n, n_buckets, people, offsets)) return rc;
    DeviceScope ds(device);
    HIP_TRY(ds.err);
    hipStream_t st = (hipStream_t)stream;
    TableScratch ts;
    if (table_scratch_alloc(ts, n) != 0) {
        table_scratch_free(ts);
        return fail(PT_ENOMEM, "radix scratch for %lld pairs", (long long)n);
    }
    hipError_t e = build_table(ts, const_cast<uint32_t*>(keys), nullptr, n, nullptr, n_buckets, ID_IDENTITY,
                               (int32_t)n, people, offsets, nullptr, st, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(st);  // the scratch is freed below
    table_scratch_free(ts);
    if (e != hipSuccess) return fail(PT_EDEVICE, "location table: %s", hipGetErrorString(e));
    return PT_OK;
}
