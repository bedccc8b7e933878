// pt_sort.hip — stable location tables: LSD radix sort of implicit-id pairs
// with wave64 match ranking and an LDS-staged scatter, then bucket offsets.
//
// Replaces cub::DeviceRadixSort::SortPairs + thrust::lower_bound
// (PandemicSim.cu:1212-1248, :1316-1354, :3835-3862).  Keys are dense bucket
// ids, so only ceil(log2 K) bits are sorted (3 passes of <= 8 bits at 50M
// agents instead of the reference's 7 passes over all 32 bits).
#include "pt_sort.h"

#include <algorithm>

namespace pt {

namespace {

#ifndef PT_SORT_THREADS
#define PT_SORT_THREADS 512
#endif
constexpr int kThreads = PT_SORT_THREADS;
constexpr int kWaves = kThreads / 64;
#ifndef PT_SORT_ITEMS
#define PT_SORT_ITEMS 8
#endif
constexpr int kItems = PT_SORT_ITEMS;  // pairs per thread per tile
constexpr int kTile = kThreads * kItems;  // pairs per tile
#ifndef PT_SORT_BLOCKS
#define PT_SORT_BLOCKS 1024
#endif
constexpr int kMaxSortBlocks = PT_SORT_BLOCKS;
constexpr int kMaxBitsSupported = 11;
constexpr int kScanSeg = 8192;  // elements per scan block
constexpr int kScanItems = kScanSeg / kThreads;

int g_max_bits = 8;  // widest digit per pass (see set_radix_max_bits)

__device__ __forceinline__ int32_t implicit_id(int64_t j, int mode, int32_t N) {
    if (mode == ID_HOUR_MAJOR) return (int32_t)(j < N ? j : j - N);
    if (mode == ID_DIV3) return (int32_t)(j / 3);
    return (int32_t)j;
}

__device__ __forceinline__ bool is_done(const int32_t* done) { return done && *done; }

// The block's slice of the n pairs (n from the device when n_dev is set).
__device__ __forceinline__ void block_range(int64_t n_host, const uint32_t* n_dev, int64_t& begin, int64_t& end) {
    const int64_t n = n_dev ? (int64_t)*n_dev : n_host;
    const int64_t tiles = (n + kTile - 1) / kTile;
    const int64_t tpb = (tiles + gridDim.x - 1) / gridDim.x;
    begin = (int64_t)blockIdx.x * tpb * kTile;
    end = min(n, begin + tpb * kTile);
}

// Inclusive scan across a wave64.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// Exclusive scan of one value per thread over the block (s_w: kWaves words).
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_w, uint32_t* total) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t inc = wave_incl_scan(v, lane);
    if (lane == 63) s_w[wave] = inc;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        const uint32_t x = s_w[w];
        if (w < wave) wbase += x;
        tot += x;
    }
    __syncthreads();
    if (total) *total = tot;
    return wbase + inc - v;
}

// Per-block digit histogram of the block's chunk.
// Keys equal to drop are left out (not counted, not written by the
// matching downsweep).  hist[nblk << BITS] is set to 0 so that the scan of
// nblk * RADIX + 1 entries leaves the pass's total there.
template <int BITS, bool MATCH>
__global__ __launch_bounds__(kThreads) void k_upsweep(const uint32_t* __restrict__ keys, int64_t n_host,
                                                      const uint32_t* __restrict__ n_dev, int shift,
                                                      uint32_t* __restrict__ hist, int nblk, const int32_t* done,
                                                      uint32_t drop) {
    if (is_done(done)) return;
    constexpr int RADIX = 1 << BITS;
    if (blockIdx.x == 0 && threadIdx.x == 0) hist[(int64_t)nblk << BITS] = 0;
    __shared__ uint32_t s_h[kWaves][RADIX];
    const int tid = threadIdx.x, wave = tid >> 6;
    for (int i = tid; i < kWaves * RADIX; i += kThreads) (&s_h[0][0])[i] = 0;
    __syncthreads();
    int64_t begin, end;
    block_range(n_host, n_dev, begin, end);
    const uint4* k4 = reinterpret_cast<const uint4*>(keys);
    if (MATCH) {
        // Keys with long runs of one digit (e.g. grouped by simulation): when
        // every active lane of the wave holds the same digit, one lane adds
        // the count; otherwise each lane adds its own.
        for (int64_t q = begin / 4 + tid; q * 4 < end; q += kThreads) {
            uint32_t kk[4];
            if (q * 4 + 3 < end) {
                const uint4 v = k4[q];
                kk[0] = v.x;
                kk[1] = v.y;
                kk[2] = v.z;
                kk[3] = v.w;
            } else {
#pragma unroll
                for (int c = 0; c < 4; ++c) kk[c] = q * 4 + c < end ? keys[q * 4 + c] : 0u;
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const bool valid = q * 4 + c < end && kk[c] != drop;
                const uint32_t d = (kk[c] >> shift) & (RADIX - 1);
                const uint64_t act = __ballot(valid);
                const int first = act ? __ffsll((unsigned long long)act) - 1 : 0;  // lowest valid lane
                const uint32_t d0 = (uint32_t)__shfl((int)d, first, 64);
                const uint64_t same = __ballot(valid && d == d0);
                if (same == act) {
                    if (valid && (threadIdx.x & 63) == first) atomicAdd(&s_h[wave][d0], (uint32_t)__popcll(act));
                } else if (valid) {
                    atomicAdd(&s_h[wave][d], 1u);
                }
            }
        }
    } else {
        for (int64_t q = begin / 4 + tid; q * 4 < end; q += kThreads) {
            if (q * 4 + 3 < end) {
                const uint4 v = k4[q];
                if (v.x != drop) atomicAdd(&s_h[wave][(v.x >> shift) & (RADIX - 1)], 1u);
                if (v.y != drop) atomicAdd(&s_h[wave][(v.y >> shift) & (RADIX - 1)], 1u);
                if (v.z != drop) atomicAdd(&s_h[wave][(v.z >> shift) & (RADIX - 1)], 1u);
                if (v.w != drop) atomicAdd(&s_h[wave][(v.w >> shift) & (RADIX - 1)], 1u);
            } else {
                for (int64_t j = q * 4; j < end; ++j)
                    if (keys[j] != drop) atomicAdd(&s_h[wave][(keys[j] >> shift) & (RADIX - 1)], 1u);
            }
        }
    }
    __syncthreads();
    for (int d = tid; d < RADIX; d += kThreads) {
        uint32_t t = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) t += s_h[w][d];
        hist[(int64_t)d * nblk + blockIdx.x] = t;
    }
}

// Stable scatter of the block's chunk by one digit.  Each tile of 4096 pairs
// is ranked per wave with ballots (match-any over the digit bits): the lowest
// lane of each digit group adds the group's size to the wave's LDS counter
// with a returning atomic, so the 8 rounds of a wave issue back to back and
// LDS order gives each round the count of the rounds before it.  The tile is
// then staged in LDS in digit order and written out as runs per digit, while
// the next tile's pairs are already in flight.
template <int BITS>
__global__ __launch_bounds__(kThreads) void k_downsweep(const uint32_t* __restrict__ keys_in,
                                                        const int32_t* __restrict__ vals_in, int id_mode, int32_t N,
                                                        uint32_t* __restrict__ keys_out, int32_t* __restrict__ vals_out,
                                                        const uint32_t* __restrict__ hist, int64_t n_host,
                                                        const uint32_t* __restrict__ n_dev, int shift, int nblk,
                                                        const int32_t* done, uint32_t drop,
                                                        uint32_t* __restrict__ n_out) {
    if (is_done(done)) return;
    constexpr int RADIX = 1 << BITS;
    // The pass's total (pairs kept) for the passes after it.
    if (n_out && blockIdx.x == 0 && threadIdx.x == 0) *n_out = hist[(int64_t)nblk << BITS];
    constexpr int DPT = RADIX > kThreads ? RADIX / kThreads : 1;  // digits per thread in the digit scan
    constexpr uint32_t MASK = RADIX - 1;
    __shared__ uint32_t s_keys[kTile];
    __shared__ int32_t s_vals[kTile];
    __shared__ uint32_t s_wcnt[kWaves][RADIX];
    __shared__ uint32_t s_start[RADIX];
    __shared__ uint32_t s_tot[RADIX];
    __shared__ uint32_t s_base[RADIX];
    __shared__ uint32_t s_delta[RADIX];
    __shared__ uint32_t s_w[kWaves];
    __shared__ uint32_t s_kept;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int64_t begin, end;
    block_range(n_host, n_dev, begin, end);
    for (int d = tid; d < RADIX; d += kThreads) s_base[d] = hist[(int64_t)d * nblk + blockIdx.x];
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    uint32_t k[kItems];
    int32_t v[kItems];
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const int64_t j = begin + wave * (kItems * 64) + r * 64 + lane;
        const bool valid = j < end;
        k[r] = valid ? keys_in[j] : 0u;
        v[r] = vals_in ? (valid ? vals_in[j] : 0) : implicit_id(j, id_mode, N);
    }

    for (int64_t t0 = begin; t0 < end; t0 += kTile) {
        const int tile_n = (int)min((int64_t)kTile, end - t0);
        for (int i = tid; i < kWaves * RADIX; i += kThreads) (&s_wcnt[0][0])[i] = 0;
        __syncthreads();
        // Next tile in flight while this one is ranked and written.
        uint32_t nk[kItems];
        int32_t nv[kItems];
        const int64_t t1 = t0 + kTile;
#pragma unroll
        for (int r = 0; r < kItems; ++r) {
            const int64_t j = t1 + wave * (kItems * 64) + r * 64 + lane;
            const bool valid = j < end;
            nk[r] = valid ? keys_in[j] : 0u;
            nv[r] = vals_in ? (valid ? vals_in[j] : 0) : implicit_id(j, id_mode, N);
        }
        uint32_t old[kItems], info[kItems];
#pragma unroll
        for (int r = 0; r < kItems; ++r) {
            const int e = wave * (kItems * 64) + r * 64 + lane;
            const bool valid = e < tile_n && k[r] != drop;
            const uint32_t d = (k[r] >> shift) & MASK;
            // Lanes whose digit differs from this lane's in some bit: per
            // bit, (ballot of the bit) XOR (this lane's bit, sign-extended).
            uint64_t peers = __ballot(valid);
            {
                uint32_t xlo = 0, xhi = 0;
#pragma unroll
                for (int b = 0; b < BITS; ++b) {
                    const int32_t ext = ((int32_t)(d << (31 - b))) >> 31;
                    const uint64_t m = __ballot(ext != 0);
                    xlo |= (uint32_t)m ^ (uint32_t)ext;
                    xhi |= (uint32_t)(m >> 32) ^ (uint32_t)ext;
                }
                peers &= ~(((uint64_t)xhi << 32) | xlo);
            }
            const uint32_t rk = (uint32_t)__popcll(peers & below);
            const uint32_t leader = peers ? (uint32_t)(__ffsll((unsigned long long)peers) - 1) : 0u;
            uint32_t o = 0;
            if (valid && rk == 0) o = atomicAdd(&s_wcnt[wave][d], (uint32_t)__popcll(peers));
            old[r] = o;
            info[r] = rk | (leader << 8);
        }
        uint32_t rank[kItems];
#pragma unroll
        for (int r = 0; r < kItems; ++r) rank[r] = __shfl(old[r], (int)(info[r] >> 8), 64) + (info[r] & 0xFF);
        __syncthreads();
        for (int d = tid; d < RADIX; d += kThreads) {
            uint32_t run = 0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                const uint32_t c = s_wcnt[w][d];
                s_wcnt[w][d] = run;
                run += c;
            }
            s_tot[d] = run;
        }
        __syncthreads();
        {
            uint32_t loc[DPT];
            uint32_t sum = 0;
#pragma unroll
            for (int j = 0; j < DPT; ++j) {
                const int d = tid * DPT + j;
                loc[j] = d < RADIX ? s_tot[d] : 0u;
                sum += loc[j];
            }
            uint32_t kept;
            uint32_t run = block_excl_scan(sum, s_w, &kept);
            if (tid == 0) s_kept = kept;
#pragma unroll
            for (int j = 0; j < DPT; ++j) {
                const int d = tid * DPT + j;
                if (d < RADIX) {
                    s_start[d] = run;
                    s_delta[d] = s_base[d] - run;  // staged pair i goes to s_delta[d] + i
                }
                run += loc[j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kItems; ++r) {
            const int e = wave * (kItems * 64) + r * 64 + lane;
            if (e < tile_n && k[r] != drop) {
                const uint32_t d = (k[r] >> shift) & MASK;
                const uint32_t pos = s_start[d] + s_wcnt[wave][d] + rank[r];
                s_keys[pos] = k[r];
                s_vals[pos] = v[r];
            }
        }
        __syncthreads();
        {
            const int kept_n = (int)s_kept;
            uint32_t kk[kItems], gg[kItems];
            int32_t vv[kItems];
#pragma unroll
            for (int j = 0; j < kItems; ++j) {
                const int i = tid + j * kThreads;
                kk[j] = i < kept_n ? s_keys[i] : 0u;
                vv[j] = i < kept_n ? s_vals[i] : 0;
            }
#pragma unroll
            for (int j = 0; j < kItems; ++j) gg[j] = s_delta[(kk[j] >> shift) & MASK] + (uint32_t)(tid + j * kThreads);
#pragma unroll
            for (int j = 0; j < kItems; ++j) {
                if (tid + j * kThreads < kept_n) {
                    keys_out[gg[j]] = kk[j];
                    vals_out[gg[j]] = vv[j];
                }
            }
        }
        __syncthreads();
        for (int d = tid; d < RADIX; d += kThreads) s_base[d] += s_tot[d];
#pragma unroll
        for (int r = 0; r < kItems; ++r) {
            k[r] = nk[r];
            v[r] = nv[r];
        }
    }
}

// Scan, phase 1: per-segment sums.
__global__ __launch_bounds__(kThreads) void k_scan_reduce(const uint32_t* __restrict__ in, int64_t n,
                                                          uint32_t* __restrict__ part, const int32_t* done) {
    if (is_done(done)) return;
    __shared__ uint32_t s_w[kWaves];
    const int64_t base = (int64_t)blockIdx.x * kScanSeg;
    uint32_t sum = 0;
    for (int i = threadIdx.x; i < kScanSeg; i += kThreads) {
        const int64_t j = base + i;
        if (j < n) sum += in[j];
    }
    uint32_t tot;
    block_excl_scan(sum, s_w, &tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// Scan, phase 2: exclusive scan of the segment sums (one block).
__global__ __launch_bounds__(kThreads) void k_scan_parts(uint32_t* part, int nparts, const int32_t* done) {
    if (is_done(done)) return;
    __shared__ uint32_t s_w[kWaves];
    uint32_t carry = 0;
    for (int b0 = 0; b0 < nparts; b0 += kThreads) {
        const int i = b0 + threadIdx.x;
        const uint32_t v = i < nparts ? part[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan(v, s_w, &tot);
        if (i < nparts) part[i] = carry + ex;
        carry += tot;
    }
}

// Scan, phase 3: exclusive scan inside each segment plus its offset.
__global__ __launch_bounds__(kThreads) void k_scan_down(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                        int64_t n, const uint32_t* __restrict__ part,
                                                        const int32_t* done) {
    if (is_done(done)) return;
    __shared__ uint32_t s_w[kWaves];
    const int64_t base = (int64_t)blockIdx.x * kScanSeg + (int64_t)threadIdx.x * kScanItems;
    uint32_t x[kScanItems];
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const int64_t j = base + i;
        x[i] = j < n ? in[j] : 0u;
        sum += x[i];
    }
    uint32_t run = block_excl_scan(sum, s_w, nullptr) + part[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const int64_t j = base + i;
        if (j < n) out[j] = run;
        run += x[i];
    }
}

// offsets[q] = first position whose key is >= q (thrust::lower_bound),
// one binary search per bucket q in [0, K].
__global__ void k_offsets(const uint32_t* __restrict__ keys, int64_t n_host, const uint32_t* __restrict__ n_dev,
                          int32_t K, int32_t* __restrict__ offsets, const int32_t* done) {
    if (is_done(done)) return;
    const int64_t n = n_dev ? (int64_t)*n_dev : n_host;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q <= K; q += stride) {
        int64_t lo = 0, hi = n;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if ((int64_t)keys[mid] < q)
                lo = mid + 1;
            else
                hi = mid;
        }
        offsets[q] = (int32_t)lo;
    }
}

template <int BITS>
hipError_t radix_pass(TableScratch& s, const uint32_t* kin, const int32_t* vin, int id_mode, int32_t N, uint32_t* kout,
                      int32_t* vout, int64_t n, const uint32_t* n_dev, int shift, int nblk, const int32_t* done,
                      hipStream_t st, LaunchTimer* tm, uint32_t drop, uint32_t* n_out) {
    if (s.grouped_keys)
        launch(tm, KC_SORT_UPSWEEP, k_upsweep<BITS, true>, dim3(nblk), dim3(kThreads), st, kin, n, n_dev, shift,
               s.hist, nblk, done, drop);
    else
        launch(tm, KC_SORT_UPSWEEP, k_upsweep<BITS, false>, dim3(nblk), dim3(kThreads), st, kin, n, n_dev, shift,
               s.hist, nblk, done, drop);
    hipError_t e =
        exclusive_scan_u32(s.hist, s.hist, ((int64_t)nblk << BITS) + 1, s.scan_part, done, st, tm, KC_SORT_SCAN);
    if (e != hipSuccess) return e;
    launch(tm, KC_SORT_DOWNSWEEP, k_downsweep<BITS>, dim3(nblk), dim3(kThreads), st, kin, vin, id_mode, N, kout, vout,
           (const uint32_t*)s.hist, n, n_dev, shift, nblk, done, drop, n_out);
    return hipGetLastError();
}

hipError_t radix_pass_bits(int bits, TableScratch& s, const uint32_t* kin, const int32_t* vin, int id_mode, int32_t N,
                           uint32_t* kout, int32_t* vout, int64_t n, const uint32_t* n_dev, int shift, int nblk,
                           const int32_t* done, hipStream_t st, LaunchTimer* tm, uint32_t drop, uint32_t* n_out) {
#define PT_PASS(B) \
    case B:        \
        return radix_pass<B>(s, kin, vin, id_mode, N, kout, vout, n, n_dev, shift, nblk, done, st, tm, drop, n_out);
    switch (bits) {
        PT_PASS(1) PT_PASS(2) PT_PASS(3) PT_PASS(4) PT_PASS(5) PT_PASS(6)
        PT_PASS(7) PT_PASS(8) PT_PASS(9) PT_PASS(10) PT_PASS(11)
    }
#undef PT_PASS
    return hipErrorInvalidValue;
}

int key_bits(int32_t K) {
    int bits = 0;
    while (((int64_t)1 << bits) < K) bits++;
    return bits == 0 ? 1 : bits;
}

}  // namespace

void set_radix_max_bits(int bits) { g_max_bits = std::min(std::max(bits, 4), kMaxBitsSupported); }

int radix_passes(int32_t K) { return (key_bits(K) + g_max_bits - 1) / g_max_bits; }

int table_scratch_alloc(TableScratch& s, int64_t n_cap) {
    table_scratch_free(s);
    const int64_t n = std::max<int64_t>(n_cap, 1);
    const int64_t hist_n = ((int64_t)1 << kMaxBitsSupported) * kMaxSortBlocks + 1;
    if (hipMalloc(&s.keys_b, n * sizeof(uint32_t)) != hipSuccess) return -1;
    if (hipMalloc(&s.vals_a, n * sizeof(int32_t)) != hipSuccess) return -1;
    if (hipMalloc(&s.vals_b, n * sizeof(int32_t)) != hipSuccess) return -1;
    if (hipMalloc(&s.hist, (size_t)hist_n * sizeof(uint32_t)) != hipSuccess) return -1;
    if (hipMalloc(&s.n_kept, sizeof(uint32_t)) != hipSuccess) return -1;
    // Partial sums for scans of up to max(n_cap, hist_n) values.
    const int64_t parts = (std::max<int64_t>(n, hist_n) + kScanSeg - 1) / kScanSeg + 2;
    if (hipMalloc(&s.scan_part, (size_t)parts * sizeof(uint32_t)) != hipSuccess) return -1;
    s.n_cap = n_cap;
    return 0;
}

void table_scratch_free(TableScratch& s) {
    if (s.keys_b) (void)hipFree(s.keys_b);
    if (s.vals_a) (void)hipFree(s.vals_a);
    if (s.vals_b) (void)hipFree(s.vals_b);
    if (s.hist) (void)hipFree(s.hist);
    if (s.n_kept) (void)hipFree(s.n_kept);
    if (s.scan_part) (void)hipFree(s.scan_part);
    s = TableScratch();
}

hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* part, const int32_t* done,
                              hipStream_t stream, LaunchTimer* tm, int cls) {
    if (n <= 0) return hipSuccess;
    const int nparts = (int)((n + kScanSeg - 1) / kScanSeg);
    launch(tm, cls, k_scan_reduce, dim3(nparts), dim3(kThreads), stream, in, n, part, done);
    launch(tm, cls, k_scan_parts, dim3(1), dim3(kThreads), stream, part, nparts, done);
    launch(tm, cls, k_scan_down, dim3(nparts), dim3(kThreads), stream, in, out, n, (const uint32_t*)part, done);
    return hipGetLastError();
}

hipError_t build_table(TableScratch& s, uint32_t* keys, const int32_t* vals, int64_t n, const uint32_t* n_dev,
                       int32_t K, IdMode mode, int32_t N, int32_t* people_out, int32_t* offsets_out,
                       const int32_t* done, hipStream_t stream, LaunchTimer* tm, uint32_t drop_key) {
    if (n > s.n_cap || K < 1 || n >= (int64_t)1 << 32) return hipErrorInvalidValue;
    if (drop_key != kNoDrop && (n_dev || drop_key < (uint32_t)K - 1u || drop_key >= (uint32_t)K))
        return hipErrorInvalidValue;  // only the last bucket of a host-sized table can be dropped
    const int bits = key_bits(K);
    const int passes = radix_passes(K);
    // Split the bits as evenly as possible over the passes.
    int pass_bits[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int p = 0; p < passes; ++p) pass_bits[p] = bits / passes + (p < bits % passes ? 1 : 0);
    uint32_t* kin = keys;
    if (n > 0) {
        // n is the capacity when n_dev carries the actual count.
        const int64_t tiles = (n + kTile - 1) / kTile;
        const int nblk = (int)std::min<int64_t>(tiles, kMaxSortBlocks);
        uint32_t* kout = s.keys_b;
        const int32_t* vin = vals;
        int shift = 0;
        for (int p = 0; p < passes; ++p) {
            int32_t* vout = (p == passes - 1) ? people_out : ((p & 1) ? s.vals_b : s.vals_a);
            const bool dropping = p == 0 && drop_key != kNoDrop;
            hipError_t e = radix_pass_bits(pass_bits[p], s, kin, vin, mode, N, kout, vout, n, n_dev, shift, nblk, done,
                                           stream, tm, dropping ? drop_key : kNoDrop, dropping ? s.n_kept : nullptr);
            if (e != hipSuccess) return e;
            // After a first pass that dropped keys the remaining passes
            // sort the pairs it kept (count on the device).
            if (dropping) n_dev = s.n_kept;
            shift += pass_bits[p];
            vin = vout;
            std::swap(kin, kout);
        }
    }
    // kin now holds the sorted keys (the caller's buffer or keys_b).
    const int oblocks = (int)std::min<int64_t>(((int64_t)K + kThreads) / kThreads, 4096);
    launch(tm, KC_OFFSETS, k_offsets, dim3(oblocks), dim3(kThreads), stream, (const uint32_t*)kin, n, n_dev, K,
           offsets_out, done);
    return hipGetLastError();
}

}  // namespace pt
