// Single-pass exclusive scan with decoupled look-back (Merrill & Garland,
// "Single-pass Parallel Prefix Scan with Decoupled Look-back", NVIDIA TR
// NVR-2016-002), written for wave64 / 256-thread workgroups.
//
// Tiles are taken in order from an atomic counter, so a tile only ever waits
// for tiles that started before it; tile 0 publishes its inclusive prefix
// without waiting, so every wait ends.  A spin bound turns a (never expected)
// stall into an error flag instead of a hang.  Status words: bits 62-63 flag
// (0 = not yet published, 1 = tile aggregate, 2 = inclusive prefix), bits 0-61
// value.  Values are sums of small counts (two 31-bit counters may be packed
// in one word); integer sums make the result independent of timing.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace zmpm {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;  // 4096
constexpr uint64_t kScanValueMask = (1ull << 62) - 1;
constexpr uint64_t kScanFlagAgg = 1ull << 62;
constexpr uint64_t kScanFlagPrefix = 2ull << 62;

struct ScanState {
    uint64_t *status;      // one word per tile, zeroed before the launch
    uint32_t *tile_ctr;    // zeroed before the launch
    uint32_t *stall_flag;  // set if a look-back exceeded its spin bound
};

// Exclusive prefix of this tile in the whole sequence, computed by one whole
// wave (all 64 lanes call it; every lane gets the result).  `agg` is the
// tile's reduction.  Lane 0 publishes the aggregate first, then the wave looks
// back over a window of 64 predecessors at once: it waits until every
// predecessor up to the nearest one that holds an inclusive prefix has
// published, sums them with one wave reduction, moves the window back if no
// inclusive prefix was in it, and finally publishes this tile's inclusive
// prefix.  Status words are read with relaxed agent-scope atomic loads (they
// carry their own value; nothing else is read through them).
__device__ __forceinline__ uint64_t lookback_wave(const ScanState &S, uint32_t tile, uint64_t agg) {
    const int lane = threadIdx.x & 63;
    if (lane == 0)
        __hip_atomic_store(&S.status[tile], (tile == 0 ? kScanFlagPrefix : kScanFlagAgg) | (agg & kScanValueMask),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint64_t excl = 0;
    if (tile > 0) {
        int64_t end = (int64_t)tile;
        uint32_t spins = 0;
        while (true) {
            const int64_t j = end - 1 - lane;
            const uint64_t w = j >= 0 ? __hip_atomic_load(&S.status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                      : kScanFlagPrefix;
            const uint64_t flag = w & ~kScanValueMask;
            const uint64_t pref = __ballot(flag == kScanFlagPrefix);
            const uint64_t none = __ballot(flag == 0);
            const int first = pref ? __ffsll((long long)pref) - 1 : 64;
            const uint64_t need = first >= 63 ? ~0ull : ((2ull << first) - 1ull);
            if (none & need) {
                if (++spins > (1u << 24)) {
                    if (lane == 0) atomicOr(S.stall_flag, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            uint64_t v = lane <= first ? (w & kScanValueMask) : 0ull;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            excl += v;
            if (first < 64) break;
            end -= 64;
        }
        if (lane == 0)
            __hip_atomic_store(&S.status[tile], kScanFlagPrefix | ((excl + agg) & kScanValueMask), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
    return excl;
}

// Exclusive scan of one 64-bit value per thread across the 256-thread block.
__device__ __forceinline__ uint64_t block_excl_scan_u64(uint64_t v, uint64_t *s_wave, uint64_t *total) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    uint64_t inc = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t y = __shfl_up(inc, off, 64);
        if (lane >= off) inc += y;
    }
    if (lane == 63) s_wave[w] = inc;
    __syncthreads();
    uint64_t pre = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint64_t s = s_wave[k];
        pre += (k < w) ? s : 0ull;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}

}  // namespace zmpm
