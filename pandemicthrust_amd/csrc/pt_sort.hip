// pt_sort.hip — stable location tables: LSD radix sort of implicit-id pairs
// with wave64 match ranking and an LDS-staged scatter, then bucket offsets.
//
// Replaces cub::DeviceRadixSort::SortPairs + thrust::lower_bound
// (PandemicSim.cu:1212-1248, :1316-1354, :3835-3862).  Keys are dense bucket
// ids, so only ceil(log2 K) bits are sorted (3 passes of <= 8 bits at 50M
// agents instead of the reference's 7 passes over all 32 bits).
//
// Data between passes: the first pass reads the caller's keys (ids explicit
// or implied by the position) and writes interleaved (key, id) pairs, one
// 8-byte store per pair; middle passes read and write pairs; the last pass
// writes the ids (the table) and, beside them, each key's bits below the
// last digit in 1 or 2 bytes: a bucket's offset is a binary search for its
// low bits inside its last digit's range, which the pass's scanned
// histogram gives (segmented builds keep whole local keys).  Passes of at
// least TableScratch::wide_bits bits (8 by default) run 1024-thread blocks
// on 8192-pair tiles (digit runs of ~32 pairs), the others 512-thread
// blocks on 4096-pair tiles.
#include "pt_sort.h"

#include <algorithm>

namespace pt {

namespace {

#ifndef PT_SORT_ITEMS
#define PT_SORT_ITEMS 8
#endif
constexpr int kItems = PT_SORT_ITEMS;  // pairs per thread per tile
constexpr int kScanThreads = 512;
constexpr int kScanSeg = 8192;  // elements per scan block
constexpr int kScanItems = kScanSeg / kScanThreads;
constexpr int kMaxBitsSupported = 11;
constexpr int kHistBlocks = 1024;  // most blocks of any pass

// Blocks of a pass with T threads: 1024 x 512 threads or 256 x 1024.
template <int T>
constexpr int max_blocks() {
    return T >= 1024 ? 256 : 1024;
}


__device__ __forceinline__ int32_t implicit_id(int64_t j, int mode, int32_t N) {
    if (mode == ID_HOUR_MAJOR) return (int32_t)(j < N ? j : j - N);
    if (mode == ID_DIV3) return (int32_t)(j / 3);
    return (int32_t)j;
}

// The same without branches (selects only), for the downsweep's unrolled
// loads: a branch there makes the compiler wait for the loads in flight.
// j / 3 as (j * 0xAAAAAAAB) >> 33, exact for every j < 2^32 (pair
// positions are below 2^32).
__device__ __forceinline__ int32_t implicit_id_select(int64_t j, int mode, int32_t N) {
    const int32_t hm = (int32_t)(j < N ? j : j - N);
    const int32_t d3 = (int32_t)(((uint64_t)(uint32_t)j * 0xAAAAAAABull) >> 33);
    return mode == ID_DIV3 ? d3 : (mode == ID_HOUR_MAJOR ? hm : (int32_t)j);
}

__device__ __forceinline__ bool is_done(const int32_t* done) { return done && *done; }

// Downsweep input: interleaved pairs, keys with ids implied by the position,
// keys with explicit ids, or (IN_DUAL, a first pass with an AltInput) keys
// with implied ids unless the device flag picks the other keys with
// explicit ids.
enum DownIn : int { IN_PAIRS = 0, IN_KEYS = 1, IN_KEYS_VALS = 2, IN_DUAL = 3 };

// Downsweep output: interleaved pairs (passes before the last), or the ids
// plus the low 8 / 16 / 32 bits of each key (byte, short or word array), or
// the ids alone.
enum DownOut : int { OUT_PAIRS = 0, OUT_IDS = 1, OUT_K8 = 2, OUT_K16 = 3, OUT_K32 = 4 };

// The first pass of a table with an AltInput reads the other input today.
__device__ __forceinline__ bool alt_today(const AltInput& a) { return a.flag && *a.flag; }

// The block's slice of the n pairs (n from the device when n_dev is set),
// whole tiles of T * kItems pairs.
template <int T>
__device__ __forceinline__ void block_range(int64_t n_host, const uint32_t* n_dev, int64_t& begin, int64_t& end) {
    constexpr int64_t tile = (int64_t)T * kItems;
    const int64_t n = n_dev ? (int64_t)*n_dev : n_host;
    const int64_t tiles = (n + tile - 1) / tile;
    const int64_t tpb = (tiles + gridDim.x - 1) / gridDim.x;
    begin = (int64_t)blockIdx.x * tpb * tile;
    end = min(n, begin + tpb * tile);
}

// Segmented passes: the block's slice of its segment (whole tiles, starting
// at a multiple of 4), the segment's alignment gap (counted by its last block
// in the top digit, so the scan leaves it after the segment's pairs), and its
// histogram column (hist[hbase + d * hstride]: segment-major, so one
// exclusive scan gives every segment its in-place offsets).
template <int T, int RADIX>
__device__ __forceinline__ void seg_range(const SegLayout& g, int64_t& begin, int64_t& end, uint32_t& gap,
                                          int64_t& hbase, int& hstride) {
    constexpr int64_t tile = (int64_t)T * kItems;
    const int r = g.blk_seg[blockIdx.x];
    const int j = (int)blockIdx.x - g.seg_blk0[r];
    const int nb = g.seg_nblk[r];
    const int64_t S = g.seg_start[r], cnt = g.seg_count[r];
    const int64_t tiles = (cnt + tile - 1) / tile;
    const int64_t tpb = (tiles + nb - 1) / nb;
    begin = S + (int64_t)j * tpb * tile;
    end = min(S + cnt, begin + tpb * tile);
    if (begin >= end) begin = end = 0;  // no pairs (keeps the vector loads of the upsweep aligned)
    gap = j == nb - 1 ? (uint32_t)(((cnt + 3) & ~(int64_t)3) - cnt) : 0u;
    hbase = (int64_t)RADIX * g.seg_blk0[r] + j;
    hstride = nb;
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

// Exclusive scan of one value per thread over a block of T threads (s_w:
// T / 64 words).
template <int T>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_w, uint32_t* total) {
    constexpr int W = T / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t inc = wave_incl_scan(v, lane);
    if (lane == 63) s_w[wave] = inc;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const uint32_t x = s_w[w];
        if (w < wave) wbase += x;
        tot += x;
    }
    __syncthreads();
    if (total) *total = tot;
    return wbase + inc - v;
}

// Per-block digit histogram of the block's chunk, from a key array (PIN
// false, 4 keys per 16-byte load) or from interleaved pairs (PIN true, 2
// pairs per load).  Keys equal to drop are left out (not counted, not
// written by the matching downsweep).  hist[nblk << BITS] is set to 0 so
// that the scan of nblk * RADIX + 1 entries leaves the pass's total there.
template <int BITS, bool MATCH, bool PIN, int T>
__global__ __launch_bounds__(T) void k_upsweep(const uint32_t* __restrict__ keys, const uint2* __restrict__ pairs,
                                               int64_t n_host, const uint32_t* __restrict__ n_dev, int shift,
                                               uint32_t* __restrict__ hist, int nblk, const int32_t* done,
                                               uint32_t drop, SegLayout seg, AltInput alt) {
    if (is_done(done)) return;
    if (!PIN && alt_today(alt)) {
        keys = alt.keys;
        n_dev = alt.n_dev;
        drop = kNoDrop;
    }
    constexpr int RADIX = 1 << BITS;
    constexpr int W = T / 64;
    constexpr int G = PIN ? 2 : 4;  // keys per 16-byte load
    if (blockIdx.x == 0 && threadIdx.x == 0) hist[(int64_t)nblk << BITS] = 0;
    __shared__ uint32_t s_h[W][RADIX];
    const int tid = threadIdx.x, wave = tid >> 6;
    for (int i = tid; i < W * RADIX; i += T) (&s_h[0][0])[i] = 0;
    __syncthreads();
    int64_t begin, end, hbase = blockIdx.x;
    int hstride = nblk;
    uint32_t gap = 0;
    if (seg.blk_seg)
        seg_range<T, RADIX>(seg, begin, end, gap, hbase, hstride);
    else
        block_range<T>(n_host, n_dev, begin, end);
    const uint4* v4 = PIN ? reinterpret_cast<const uint4*>(pairs) : reinterpret_cast<const uint4*>(keys);
    // U 16-byte loads in flight per thread and iteration.
    constexpr int U = 4;
    for (int64_t q0 = begin / G + tid; q0 * G < end; q0 += (int64_t)U * T) {
        uint32_t kk[U][G];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t q = q0 + (int64_t)u * T;
            if (q * G + G - 1 < end) {
                const uint4 v = v4[q];
                if (PIN) {
                    kk[u][0] = v.x;
                    kk[u][G - 1] = v.z;
                } else {
                    kk[u][0] = v.x;
                    kk[u][1 % G] = v.y;
                    kk[u][2 % G] = v.z;
                    kk[u][3 % G] = v.w;
                }
            } else {
#pragma unroll
                for (int c = 0; c < G; ++c) {
                    const int64_t j = q * G + c;
                    kk[u][c] = j < end ? (PIN ? pairs[j].x : keys[j]) : 0u;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int c = 0; c < G; ++c) {
                const bool valid = (q0 + (int64_t)u * T) * G + c < end && kk[u][c] != drop;
                const uint32_t d = (kk[u][c] >> shift) & (RADIX - 1);
                if (MATCH) {
                    // Keys with long runs of one digit (e.g. grouped by
                    // simulation): when every active lane of the wave holds
                    // the same digit, one lane adds the count.
                    const uint64_t act = __ballot(valid);
                    const int first = act ? __ffsll((unsigned long long)act) - 1 : 0;  // lowest valid lane
                    const uint32_t d0 = (uint32_t)__shfl((int)d, first, 64);
                    const uint64_t same = __ballot(valid && d == d0);
                    if (same == act) {
                        if (valid && (threadIdx.x & 63) == first)
                            atomicAdd(&s_h[wave][d0], (uint32_t)__popcll(act));
                    } else if (valid) {
                        atomicAdd(&s_h[wave][d], 1u);
                    }
                } else if (valid) {
                    atomicAdd(&s_h[wave][d], 1u);
                }
            }
        }
    }
    __syncthreads();
    for (int d = tid; d < RADIX; d += T) {
        uint32_t t = d == RADIX - 1 ? gap : 0u;
#pragma unroll
        for (int w = 0; w < W; ++w) t += s_h[w][d];
        hist[hbase + (int64_t)d * hstride] = t;
    }
}

// Stable scatter of the block's chunk by one digit.  Each tile of T * 8
// pairs is ranked per wave with ballots (match-any over the digit bits): the
// lowest lane of each digit group adds the group's size to the wave's LDS
// counter with a returning atomic, so the 8 rounds of a wave issue back to
// back and LDS order gives each round the count of the rounds before it.
// The tile is then staged in LDS in digit order and written out as runs per
// digit, while the next tile's pairs are already in flight.  Input (IN):
// keys with ids implied by the position, keys with explicit ids, or pairs;
// each is its own instantiation, so the unrolled loads are branch-free and
// stay in flight.  Output (OUT): pairs, or ids with the keys' low 8, 16 or
// 32 bits beside them, or ids alone.
template <int BITS, int T>
struct DownSmem {
    uint2 pairs[T * kItems];
    uint32_t wcnt[T / 64][1 << BITS];
    uint32_t start[1 << BITS];
    uint32_t tot[1 << BITS];
    uint32_t base[1 << BITS];
    uint32_t delta[1 << BITS];
    uint32_t w[T / 64];
    uint32_t kept;
};

template <int BITS, int IN, int OUT, int T>
__device__ __forceinline__ void downsweep_body(DownSmem<BITS, T>& sm, const uint32_t* __restrict__ keys_in,
                                               const int32_t* __restrict__ vals_in,
                                               const uint2* __restrict__ pairs_in, int id_mode, int32_t N,
                                               uint32_t* __restrict__ keys_out, int32_t* __restrict__ vals_out,
                                               uint2* __restrict__ pairs_out, const uint32_t* __restrict__ hist,
                                               int64_t n_host, const uint32_t* __restrict__ n_dev, int shift,
                                               int nblk, uint32_t drop, uint32_t* __restrict__ n_out,
                                               const SegLayout& seg) {
    constexpr bool PIN = IN == IN_PAIRS;
    constexpr int RADIX = 1 << BITS;
    constexpr int W = T / 64;
    constexpr int TILE = T * kItems;
    // The pass's total (pairs kept) for the passes after it.
    if (n_out && blockIdx.x == 0 && threadIdx.x == 0) *n_out = hist[(int64_t)nblk << BITS];
    constexpr int DPT = RADIX > T ? RADIX / T : 1;  // digits per thread in the digit scan
    constexpr uint32_t MASK = RADIX - 1;
    uint2* s_pairs = sm.pairs;
    auto& s_wcnt = sm.wcnt;
    uint32_t* s_start = sm.start;
    uint32_t* s_tot = sm.tot;
    uint32_t* s_base = sm.base;
    uint32_t* s_delta = sm.delta;
    uint32_t* s_w = sm.w;
    uint32_t& s_kept = sm.kept;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int64_t begin, end;
    int64_t hbase = blockIdx.x;
    int hstride = nblk;
    if (seg.blk_seg) {
        uint32_t gap;
        seg_range<T, RADIX>(seg, begin, end, gap, hbase, hstride);
    } else {
        block_range<T>(n_host, n_dev, begin, end);
    }
    for (int d = tid; d < RADIX; d += T) s_base[d] = hist[hbase + (int64_t)d * hstride];
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    auto load = [&](int64_t j, bool valid, uint32_t& k, int32_t& v) {
        if (PIN) {
            const uint2 p = valid ? pairs_in[j] : make_uint2(0u, 0u);
            k = p.x;
            v = (int32_t)p.y;
        } else {
            k = valid ? keys_in[j] : 0u;
            v = IN == IN_KEYS_VALS ? (valid ? vals_in[j] : 0) : implicit_id_select(j, id_mode, N);
        }
    };

    uint32_t k[kItems];
    int32_t v[kItems];
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const int64_t j = begin + wave * (kItems * 64) + r * 64 + lane;
        load(j, j < end, k[r], v[r]);
    }

    for (int64_t t0 = begin; t0 < end; t0 += TILE) {
        const int tile_n = (int)min((int64_t)TILE, end - t0);
        for (int i = tid; i < W * RADIX; i += T) (&s_wcnt[0][0])[i] = 0;
        __syncthreads();
        // Next tile in flight while this one is ranked and written.
        uint32_t nk[kItems];
        int32_t nv[kItems];
        const int64_t t1 = t0 + TILE;
#pragma unroll
        for (int r = 0; r < kItems; ++r) {
            const int64_t j = t1 + wave * (kItems * 64) + r * 64 + lane;
            load(j, j < end, nk[r], nv[r]);
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
        for (int d = tid; d < RADIX; d += T) {
            uint32_t run = 0;
#pragma unroll
            for (int w = 0; w < W; ++w) {
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
            uint32_t run = block_excl_scan<T>(sum, s_w, &kept);
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
                s_pairs[pos] = make_uint2(k[r], (uint32_t)v[r]);
            }
        }
        __syncthreads();
        {
            const int kept_n = (int)s_kept;
            uint2 pp[kItems];
            uint32_t gg[kItems];
#pragma unroll
            for (int j = 0; j < kItems; ++j) {
                const int i = tid + j * T;
                pp[j] = i < kept_n ? s_pairs[i] : make_uint2(0u, 0u);
            }
#pragma unroll
            for (int j = 0; j < kItems; ++j) gg[j] = s_delta[(pp[j].x >> shift) & MASK] + (uint32_t)(tid + j * T);
#pragma unroll
            for (int j = 0; j < kItems; ++j) {
                if (tid + j * T < kept_n) {
                    if constexpr (OUT == OUT_PAIRS) {
                        pairs_out[gg[j]] = pp[j];
                    } else {
                        vals_out[gg[j]] = (int32_t)pp[j].y;
                        if constexpr (OUT == OUT_K8)
                            reinterpret_cast<uint8_t*>(keys_out)[gg[j]] = (uint8_t)pp[j].x;
                        else if constexpr (OUT == OUT_K16)
                            reinterpret_cast<uint16_t*>(keys_out)[gg[j]] = (uint16_t)pp[j].x;
                        else if constexpr (OUT == OUT_K32)
                            keys_out[gg[j]] = pp[j].x;
                    }
                }
            }
        }
        __syncthreads();
        for (int d = tid; d < RADIX; d += T) s_base[d] += s_tot[d];
#pragma unroll
        for (int r = 0; r < kItems; ++r) {
            k[r] = nk[r];
            v[r] = nv[r];
        }
    }
}

// Up to 8-bit digits a pass keeps 4 waves per SIMD (at most 128 VGPRs): the
// two-input first pass would otherwise take a few registers more and lose
// one of them.
template <int BITS, int IN, int OUT, int T>
__global__ __launch_bounds__(T, BITS <= 8 ? 4 : 1) void k_downsweep(const uint32_t* __restrict__ keys_in,
                                                 const int32_t* __restrict__ vals_in, const uint2* __restrict__ pairs_in,
                                                 int id_mode, int32_t N, uint32_t* __restrict__ keys_out,
                                                 int32_t* __restrict__ vals_out, uint2* __restrict__ pairs_out,
                                                 const uint32_t* __restrict__ hist, int64_t n_host,
                                                 const uint32_t* __restrict__ n_dev, int shift, int nblk,
                                                 const int32_t* done, uint32_t drop, uint32_t* __restrict__ n_out,
                                                 SegLayout seg, AltInput alt) {
    if (is_done(done)) return;
    __shared__ DownSmem<BITS, T> sm;
    if constexpr (IN == IN_DUAL) {
        if (alt_today(alt))
            downsweep_body<BITS, IN_KEYS_VALS, OUT, T>(sm, alt.keys, alt.vals, nullptr, id_mode, N, keys_out,
                                                        vals_out, pairs_out, hist, n_host, alt.n_dev, shift, nblk,
                                                        kNoDrop, n_out, seg);
        else
            downsweep_body<BITS, IN_KEYS, OUT, T>(sm, keys_in, nullptr, nullptr, id_mode, N, keys_out, vals_out,
                                                   pairs_out, hist, n_host, n_dev, shift, nblk, drop, n_out, seg);
    } else {
        downsweep_body<BITS, IN, OUT, T>(sm, keys_in, vals_in, pairs_in, id_mode, N, keys_out, vals_out, pairs_out,
                                          hist, n_host, n_dev, shift, nblk, drop, n_out, seg);
    }
}

// Scan, phase 1: per-segment sums.
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t* __restrict__ in, int64_t n,
                                                              uint32_t* __restrict__ part, const int32_t* done) {
    if (is_done(done)) return;
    __shared__ uint32_t s_w[kScanThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanSeg;
    uint32_t sum = 0;
    for (int i = threadIdx.x; i < kScanSeg; i += kScanThreads) {
        const int64_t j = base + i;
        if (j < n) sum += in[j];
    }
    uint32_t tot;
    block_excl_scan<kScanThreads>(sum, s_w, &tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// Scan, phase 2: exclusive scan of the segment sums (one block).
__global__ __launch_bounds__(kScanThreads) void k_scan_parts(uint32_t* part, int nparts, const int32_t* done) {
    if (is_done(done)) return;
    __shared__ uint32_t s_w[kScanThreads / 64];
    uint32_t carry = 0;
    for (int b0 = 0; b0 < nparts; b0 += kScanThreads) {
        const int i = b0 + threadIdx.x;
        const uint32_t v = i < nparts ? part[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan<kScanThreads>(v, s_w, &tot);
        if (i < nparts) part[i] = carry + ex;
        carry += tot;
    }
}

// Scan, phase 3: exclusive scan inside each segment plus its offset.  With
// SUM_PARTS the offset is summed here from the raw segment sums (at most
// kScanFewParts segments), so phase 2 is not launched.
constexpr int kScanFewParts = 64;

__device__ __forceinline__ uint32_t wave_total(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <bool SUM_PARTS>
__global__ __launch_bounds__(kScanThreads) void k_scan_down(const uint32_t* __restrict__ in,
                                                            uint32_t* __restrict__ out, int64_t n,
                                                            const uint32_t* __restrict__ part, const int32_t* done) {
    if (is_done(done)) return;
    __shared__ uint32_t s_w[kScanThreads / 64];
    __shared__ uint32_t s_off;
    if (SUM_PARTS) {
        if (threadIdx.x < 64) {
            uint32_t v = (int)threadIdx.x < (int)blockIdx.x ? part[threadIdx.x] : 0u;
            v = wave_total(v);
            if (threadIdx.x == 0) s_off = v;
        }
        __syncthreads();
    }
    const int64_t base = (int64_t)blockIdx.x * kScanSeg + (int64_t)threadIdx.x * kScanItems;
    uint32_t x[kScanItems];
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const int64_t j = base + i;
        x[i] = j < n ? in[j] : 0u;
        sum += x[i];
    }
    uint32_t run = block_excl_scan<kScanThreads>(sum, s_w, nullptr) + (SUM_PARTS ? s_off : part[blockIdx.x]);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const int64_t j = base + i;
        if (j < n) out[j] = run;
        run += x[i];
    }
}

// offsets[q] = first position whose key is >= q (thrust::lower_bound), for
// q in [0, K], after a last pass over digit bits [low_bits, low_bits +
// top_bits).  hist is that pass's scanned histogram (digit-major, nblk
// blocks), so the keys with top digit d = q >> low_bits occupy [hist[d *
// nblk], hist[(d + 1) * nblk]) (the total past the last digit), sorted by
// their low bits; low holds each pair's key truncated to KW bytes (KW * 8 >=
// low_bits, so the bits it keeps above low_bits are d's, equal within the
// range), searched for q truncated alike.  KW = 0: the digit is the key.
template <int KW>
__device__ __forceinline__ uint32_t low_key(const void* low, int64_t i) {
    if constexpr (KW == 1) return reinterpret_cast<const uint8_t*>(low)[i];
    else if constexpr (KW == 2) return reinterpret_cast<const uint16_t*>(low)[i];
    else return reinterpret_cast<const uint32_t*>(low)[i];
}

// Offsets of a table without pairs: every bucket empty.
__global__ void k_fill_offsets(int32_t K, int32_t* __restrict__ offsets, const int32_t* done) {
    if (is_done(done)) return;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q <= K; q += stride) offsets[q] = 0;
}

template <int KW>
__global__ void k_offsets(const void* __restrict__ low, const uint32_t* __restrict__ hist, int nblk, int top_bits,
                          int low_bits, int32_t K, int32_t* __restrict__ offsets, const int32_t* done) {
    if (is_done(done)) return;
    const int64_t radix = (int64_t)1 << top_bits;
    const int64_t total = hist[(int64_t)nblk << top_bits];
    const uint32_t mask = KW == 1 ? 0xFFu : (KW == 2 ? 0xFFFFu : 0xFFFFFFFFu);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q <= K; q += stride) {
        const int64_t d = q >> low_bits;
        if (q == K || d >= radix) {
            offsets[q] = (int32_t)total;
            continue;
        }
        int64_t a = hist[d * nblk];
        if constexpr (KW > 0) {
            int64_t b = d + 1 < radix ? (int64_t)hist[(d + 1) * nblk] : total;
            const uint32_t t = (uint32_t)q & mask;
            while (a < b) {
                const int64_t mid = (a + b) >> 1;
                if (low_key<KW>(low, mid) < t)
                    a = mid + 1;
                else
                    b = mid;
            }
        }
        offsets[q] = (int32_t)a;
    }
}

// Segmented offsets: bucket q of the global numbering belongs to the last
// segment r with seg_kbase[r] <= q; offsets[q] = the first position of
// segment r whose local key is >= q - seg_kbase[r] (its whole range when q
// lies past the segment's buckets).
__global__ void k_offsets_seg(const uint32_t* __restrict__ keys, SegLayout seg, int32_t K_total,
                              int32_t* __restrict__ offsets, const int32_t* done) {
    if (is_done(done)) return;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q <= K_total; q += stride) {
        int lo = 0, hi = seg.R;  // last r in [0, R) with seg_kbase[r] <= q
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if ((int64_t)seg.seg_kbase[mid] <= q)
                lo = mid;
            else
                hi = mid;
        }
        const int r = lo;
        const int64_t S = seg.seg_start[r], cnt = seg.seg_count[r];
        const int64_t lq = q - (int64_t)seg.seg_kbase[r];
        int64_t a = S + cnt;  // at or past the segment's last bucket: its end
        if (lq < seg.seg_K[r]) {
            int64_t b = S + cnt;
            a = S;
            while (a < b) {
                const int64_t mid = (a + b) >> 1;
                if ((int64_t)keys[mid] < lq)
                    a = mid + 1;
                else
                    b = mid;
            }
        }
        offsets[q] = (int32_t)a;
    }
}

// One radix pass.  Input: keys + vals (or implied ids) when pin is null,
// else pairs pin; output: pairs pout when set, else keys kout + ids vout.
struct PassIO {
    const uint32_t* kin;
    const int32_t* vin;
    const uint2* pin;
    uint32_t* kout;  // last pass: the keys' low kw bytes (kw 1, 2 or 4; 0: none)
    int32_t* vout;
    uint2* pout;
    int kw;
    AltInput alt;    // first pass only: the input the device may pick instead of kin
    int* nblk_out;   // set to the pass's block count when not null
};

template <int BITS, int T>
hipError_t radix_pass(TableScratch& s, const PassIO& io, int id_mode, int32_t N, int64_t n, const uint32_t* n_dev,
                      int shift, const int32_t* done, hipStream_t st, LaunchTimer* tm, uint32_t drop,
                      uint32_t* n_out, const SegLayout& seg) {
    const int64_t tiles = (n + (int64_t)T * kItems - 1) / ((int64_t)T * kItems);
    const int nblk = seg.blk_seg ? kSegBlocks : (int)std::min<int64_t>(std::max<int64_t>(tiles, 1), max_blocks<T>());
    if (io.nblk_out) *io.nblk_out = nblk;
    const bool pin = io.pin != nullptr;
    if (!io.pout && ((pin && io.kw == 0) || (!pin && io.kw != 0 && (io.kw != 4 || !io.vin || io.alt.flag))))
        return hipErrorInvalidValue;  // the output forms below are the only ones instantiated
    const bool grouped = s.grouped_keys && !seg.blk_seg;  // segment-local keys do not come in long runs
    if (pin && grouped)
        launch(tm, KC_SORT_UPSWEEP, k_upsweep<BITS, true, true, T>, dim3(nblk), dim3(T), st, io.kin, io.pin, n, n_dev,
               shift, s.hist, nblk, done, drop, seg, io.alt);
    else if (pin)
        launch(tm, KC_SORT_UPSWEEP, k_upsweep<BITS, false, true, T>, dim3(nblk), dim3(T), st, io.kin, io.pin, n, n_dev,
               shift, s.hist, nblk, done, drop, seg, io.alt);
    else if (grouped)
        launch(tm, KC_SORT_UPSWEEP, k_upsweep<BITS, true, false, T>, dim3(nblk), dim3(T), st, io.kin, io.pin, n, n_dev,
               shift, s.hist, nblk, done, drop, seg, io.alt);
    else
        launch(tm, KC_SORT_UPSWEEP, k_upsweep<BITS, false, false, T>, dim3(nblk), dim3(T), st, io.kin, io.pin, n,
               n_dev, shift, s.hist, nblk, done, drop, seg, io.alt);
    hipError_t e =
        exclusive_scan_u32(s.hist, s.hist, ((int64_t)nblk << BITS) + 1, s.scan_part, done, st, tm, KC_SORT_SCAN);
    if (e != hipSuccess) return e;
    const uint32_t* h = s.hist;
#define PT_DOWN(PI, PO)                                                                                              \
    launch(tm, KC_SORT_DOWNSWEEP, k_downsweep<BITS, PI, PO, T>, dim3(nblk), dim3(T), st, io.kin, io.vin, io.pin,      \
           id_mode, N, io.kout, io.vout, io.pout, h, n, n_dev, shift, nblk, done, drop, n_out, seg, io.alt)
    if (io.pout) {
        if (io.alt.flag)
            PT_DOWN(IN_DUAL, OUT_PAIRS);
        else if (pin)
            PT_DOWN(IN_PAIRS, OUT_PAIRS);
        else if (io.vin)
            PT_DOWN(IN_KEYS_VALS, OUT_PAIRS);
        else
            PT_DOWN(IN_KEYS, OUT_PAIRS);
    } else if (pin) {  // the last of several passes
        if (io.kw == 1)
            PT_DOWN(IN_PAIRS, OUT_K8);
        else if (io.kw == 2)
            PT_DOWN(IN_PAIRS, OUT_K16);
        else
            PT_DOWN(IN_PAIRS, OUT_K32);
    } else if (io.kw == 4) {  // one-pass segmented tables keep their local keys
        PT_DOWN(IN_KEYS_VALS, OUT_K32);
    } else {  // one-pass tables: the digit is the key
        if (io.alt.flag)
            PT_DOWN(IN_DUAL, OUT_IDS);
        else if (io.vin)
            PT_DOWN(IN_KEYS_VALS, OUT_IDS);
        else
            PT_DOWN(IN_KEYS, OUT_IDS);
    }
#undef PT_DOWN
    return hipGetLastError();
}

// 1024-thread blocks exist for digits of up to 8 bits (LDS).
hipError_t radix_pass_bits(int bits, TableScratch& s, const PassIO& io, int id_mode, int32_t N, int64_t n,
                           const uint32_t* n_dev, int shift, const int32_t* done, hipStream_t st, LaunchTimer* tm,
                           uint32_t drop, uint32_t* n_out, const SegLayout& seg) {
    // Segmented passes always run kSegBlocks blocks of 512 threads.
    const bool wide = !seg.blk_seg && bits <= 8 && bits >= s.wide_bits;
#define PT_PASS2(B)                                                                                            \
    case B:                                                                                                    \
        return wide ? radix_pass<B, 1024>(s, io, id_mode, N, n, n_dev, shift, done, st, tm, drop, n_out, seg)  \
                    : radix_pass<B, 512>(s, io, id_mode, N, n, n_dev, shift, done, st, tm, drop, n_out, seg);
#define PT_PASS1(B) \
    case B: return radix_pass<B, 512>(s, io, id_mode, N, n, n_dev, shift, done, st, tm, drop, n_out, seg);
    switch (bits) {
        PT_PASS2(1) PT_PASS2(2) PT_PASS2(3) PT_PASS2(4) PT_PASS2(5) PT_PASS2(6)
        PT_PASS2(7) PT_PASS2(8) PT_PASS1(9) PT_PASS1(10) PT_PASS1(11)
    }
#undef PT_PASS1
#undef PT_PASS2
    return hipErrorInvalidValue;
}

int key_bits(int32_t K) {
    int bits = 0;
    while (((int64_t)1 << bits) < K) bits++;
    return bits == 0 ? 1 : bits;
}

}  // namespace

int radix_passes(int32_t K, int max_bits) {
    const int mb = std::min(std::max(max_bits, 4), kMaxBitsSupported);
    return (key_bits(K) + mb - 1) / mb;
}

int table_scratch_alloc(TableScratch& s, int64_t n_cap) {
    const int wide_bits = s.wide_bits, max_bits = s.max_bits;  // launch settings, kept across reallocation
    table_scratch_free(s);
    const int64_t n = std::max<int64_t>(n_cap, 1);
    const int64_t hist_n = ((int64_t)1 << kMaxBitsSupported) * kHistBlocks + 1;
    if (hipMalloc(&s.keys_b, n * sizeof(uint32_t)) != hipSuccess) return -1;
    if (hipMalloc(&s.pairs_a, n * sizeof(uint2)) != hipSuccess) return -1;
    if (hipMalloc(&s.pairs_b, n * sizeof(uint2)) != hipSuccess) return -1;
    if (hipMalloc(&s.hist, (size_t)hist_n * sizeof(uint32_t)) != hipSuccess) return -1;
    if (hipMalloc(&s.n_kept, sizeof(uint32_t)) != hipSuccess) return -1;
    // Partial sums for scans of up to max(n_cap, hist_n) values.
    const int64_t parts = (std::max<int64_t>(n, hist_n) + kScanSeg - 1) / kScanSeg + 2;
    if (hipMalloc(&s.scan_part, (size_t)parts * sizeof(uint32_t)) != hipSuccess) return -1;
    s.n_cap = n_cap;
    s.wide_bits = wide_bits;
    s.max_bits = max_bits;
    return 0;
}

void table_scratch_free(TableScratch& s) {
    if (s.keys_b) (void)hipFree(s.keys_b);
    if (s.pairs_a) (void)hipFree(s.pairs_a);
    if (s.pairs_b) (void)hipFree(s.pairs_b);
    if (s.hist) (void)hipFree(s.hist);
    if (s.n_kept) (void)hipFree(s.n_kept);
    if (s.scan_part) (void)hipFree(s.scan_part);
    s = TableScratch();
}

hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* part, const int32_t* done,
                              hipStream_t stream, LaunchTimer* tm, int cls) {
    if (n <= 0) return hipSuccess;
    const int nparts = (int)((n + kScanSeg - 1) / kScanSeg);
    launch(tm, cls, k_scan_reduce, dim3(nparts), dim3(kScanThreads), stream, in, n, part, done);
    if (nparts <= kScanFewParts) {
        launch(tm, cls, k_scan_down<true>, dim3(nparts), dim3(kScanThreads), stream, in, out, n,
               (const uint32_t*)part, done);
    } else {
        launch(tm, cls, k_scan_parts, dim3(1), dim3(kScanThreads), stream, part, nparts, done);
        launch(tm, cls, k_scan_down<false>, dim3(nparts), dim3(kScanThreads), stream, in, out, n,
               (const uint32_t*)part, done);
    }
    return hipGetLastError();
}

hipError_t build_table(TableScratch& s, uint32_t* keys, const int32_t* vals, int64_t n, const uint32_t* n_dev,
                       int32_t K, IdMode mode, int32_t N, int32_t* people_out, int32_t* offsets_out,
                       const int32_t* done, hipStream_t stream, LaunchTimer* tm, uint32_t drop_key,
                       const AltInput* alt) {
    if (n > s.n_cap || K < 1 || n >= (int64_t)1 << 32) return hipErrorInvalidValue;
    if (drop_key != kNoDrop && (n_dev || drop_key < (uint32_t)K - 1u || drop_key >= (uint32_t)K))
        return hipErrorInvalidValue;  // only the last bucket of a host-sized table can be dropped
    if (alt && (!alt->flag || !alt->keys || !alt->vals || !alt->n_dev || vals || n_dev))
        return hipErrorInvalidValue;  // the alternative is explicit ids counted on the device; the main input is not
    const int bits = key_bits(K);
    const int passes = radix_passes(K, s.max_bits);
    // Split the bits as evenly as possible over the passes.
    int pass_bits[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int p = 0; p < passes; ++p) pass_bits[p] = bits / passes + (p < bits % passes ? 1 : 0);
    const int oblocks = (int)std::min<int64_t>(((int64_t)K + 512) / 512, 4096);
    if (n == 0) {
        launch(tm, KC_OFFSETS, k_fill_offsets, dim3(oblocks), dim3(512), stream, K, offsets_out, done);
        return hipGetLastError();
    }
    // The last pass writes the ids and, beside them, each key's bits below
    // its digit (low_bits) in the narrowest of 1, 2 or 4 bytes; the offsets
    // search those inside the digit's range.
    const int low_bits = bits - pass_bits[passes - 1];
    const int kw = low_bits == 0 ? 0 : (low_bits <= 8 ? 1 : (low_bits <= 16 ? 2 : 4));
    int last_nblk = 0;
    {
        // n is the capacity when n_dev carries the actual count.
        int shift = 0;
        for (int p = 0; p < passes; ++p) {
            const bool last = p == passes - 1;
            PassIO io{};
            if (p == 0) {
                io.kin = keys;
                io.vin = vals;
                if (alt) io.alt = *alt;
            } else {
                io.pin = (p & 1) ? s.pairs_a : s.pairs_b;
            }
            if (last) {
                io.kout = s.keys_b;
                io.vout = people_out;
                io.kw = kw;
                io.nblk_out = &last_nblk;
            } else {
                io.pout = (p & 1) ? s.pairs_b : s.pairs_a;
            }
            const bool dropping = p == 0 && drop_key != kNoDrop;
            const bool counting = p == 0 && (dropping || alt);
            hipError_t e = radix_pass_bits(pass_bits[p], s, io, mode, N, n, n_dev, shift, done, stream, tm,
                                           dropping ? drop_key : kNoDrop, counting ? s.n_kept : nullptr, SegLayout{});
            if (e != hipSuccess) return e;
            // After a first pass that dropped keys (or may have read the
            // alternative input) the remaining passes sort the pairs it
            // kept (count on the device).
            if (counting) n_dev = s.n_kept;
            shift += pass_bits[p];
        }
    }
    const int top_bits = pass_bits[passes - 1];
    const void* low = s.keys_b;
    const uint32_t* h = s.hist;
    if (kw == 0)
        launch(tm, KC_OFFSETS, k_offsets<0>, dim3(oblocks), dim3(512), stream, low, h, last_nblk, top_bits, low_bits,
               K, offsets_out, done);
    else if (kw == 1)
        launch(tm, KC_OFFSETS, k_offsets<1>, dim3(oblocks), dim3(512), stream, low, h, last_nblk, top_bits, low_bits,
               K, offsets_out, done);
    else if (kw == 2)
        launch(tm, KC_OFFSETS, k_offsets<2>, dim3(oblocks), dim3(512), stream, low, h, last_nblk, top_bits, low_bits,
               K, offsets_out, done);
    else
        launch(tm, KC_OFFSETS, k_offsets<4>, dim3(oblocks), dim3(512), stream, low, h, last_nblk, top_bits, low_bits,
               K, offsets_out, done);
    return hipGetLastError();
}

hipError_t build_table_segmented(TableScratch& s, uint32_t* keys, const int32_t* vals, int64_t n_cap,
                                 int32_t K_local, int32_t K_total, const SegLayout& seg, int32_t* people_out,
                                 int32_t* offsets_out, const int32_t* done, hipStream_t stream, LaunchTimer* tm) {
    if (n_cap > s.n_cap || K_local < 1 || K_total < 1 || !vals || !seg.blk_seg || seg.R < 1 || seg.R > kSegBlocks)
        return hipErrorInvalidValue;
    const int bits = key_bits(K_local);
    const int passes = radix_passes(K_local, s.max_bits);
    int pass_bits[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int p = 0; p < passes; ++p) pass_bits[p] = bits / passes + (p < bits % passes ? 1 : 0);
    int shift = 0;
    for (int p = 0; p < passes; ++p) {
        const bool last = p == passes - 1;
        PassIO io{};
        if (p == 0) {
            io.kin = keys;
            io.vin = vals;
        } else {
            io.pin = (p & 1) ? s.pairs_a : s.pairs_b;
        }
        if (last) {
            io.kout = s.keys_b;  // whole local keys for k_offsets_seg
            io.vout = people_out;
            io.kw = 4;
        } else {
            io.pout = (p & 1) ? s.pairs_b : s.pairs_a;
        }
        hipError_t e = radix_pass_bits(pass_bits[p], s, io, ID_IDENTITY, 0, n_cap, nullptr, shift, done, stream, tm,
                                       kNoDrop, nullptr, seg);
        if (e != hipSuccess) return e;
        shift += pass_bits[p];
    }
    const int oblocks = (int)std::min<int64_t>(((int64_t)K_total + 512) / 512, 4096);
    launch(tm, KC_OFFSETS, k_offsets_seg, dim3(oblocks), dim3(512), stream, (const uint32_t*)s.keys_b, seg, K_total,
           offsets_out, done);
    return hipGetLastError();
}

}  // namespace pt
