// Device-wide exclusive scan and stable LSD radix sort for gfx950 (wave64).
//
// The partition step of the reference (ZSPartitionForZSParticles,
// projects/CuLagrange/mpm/MPMPipeline.cu:75-84) inserts block keys into a
// hash table; its exclusive scans (MPMPipeline.cu:210) and the zpc
// radix_sort_pair used for Morton ordering (projects/CUDA/Structures.hpp:430)
// are the library calls these kernels replace.  Here particles are binned by a
// stable sort of (block key, index), so the particle order and block list are
// deterministic and equal to std::stable_sort by key.
//
// Radix sort: 8-bit digits, tiles of 2048 keys (256 threads x 8 rounds).
//   hist    per-tile digit counts, stored digit-major so one scan gives every
//           (digit, tile) output offset;
//   scan    device-wide exclusive scan (reduce / scan partials / downsweep);
//   scatter stable in-tile ranks from wave ballots (8 ballots match the digit
//           across the 64 lanes), per-wave counts in LDS for the cross-wave
//           prefix; keys are staged in LDS in digit order so each digit run is
//           written out with consecutive addresses.
#include "zmpm_internal.h"

namespace zmpm {

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kChunk = kThreads * kItems;  // 2048
constexpr int kRadix = 256;
static_assert(kChunk == kSortChunk && kRadix == kSortRadix, "radix_scratch_bound's tile and digit sizes");

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x, int lane) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    return x;
}

// Exclusive scan of one value per thread over a 256-thread block; returns the
// exclusive prefix and writes the block total to *total.
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t *s_wave, uint32_t *total) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    uint32_t inc = wave_incl_scan(v, lane);
    if (lane == 63) s_wave[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint32_t s = s_wave[k];
        pre += (k < w) ? s : 0u;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}

__global__ void __launch_bounds__(kThreads) k_fill(uint32_t *p, uint32_t value, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
        p[i] = value;
}

__device__ __forceinline__ uint32_t scan_in(const uint32_t *in, int64_t i, int shift, uint32_t mask) {
    return (in[i] >> shift) & mask;
}

__global__ void __launch_bounds__(kThreads) k_scan_reduce(const uint32_t *in, int64_t n, int shift, uint32_t mask,
                                                          uint32_t *partial) {
    __shared__ uint32_t s_wave[4];
    const int64_t base = (int64_t)blockIdx.x * kChunk + (int64_t)threadIdx.x * kItems;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        int64_t i = base + k;
        if (i < n) s += scan_in(in, i, shift, mask);
    }
    uint32_t tot;
    block_excl_scan256(s, s_wave, &tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Single workgroup: exclusive scan of the per-chunk partial sums, in place.
__global__ void __launch_bounds__(1024) k_scan_partials(uint32_t *partial, int64_t np) {
    __shared__ uint32_t s_wave[16];
    __shared__ uint32_t s_carry;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) s_carry = 0;
    __syncthreads();
    for (int64_t b = 0; b < np; b += 1024) {
        int64_t i = b + t;
        uint32_t v = i < np ? partial[i] : 0u;
        uint32_t inc = wave_incl_scan(v, lane);
        if (lane == 63) s_wave[w] = inc;
        __syncthreads();
        uint32_t pre = 0, tot = 0;
        for (int k = 0; k < 16; ++k) {
            uint32_t s = s_wave[k];
            pre += (k < w) ? s : 0u;
            tot += s;
        }
        uint32_t carry = s_carry;
        if (i < np) partial[i] = carry + pre + inc - v;
        __syncthreads();
        if (t == 0) s_carry = carry + tot;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kThreads) k_scan_down(const uint32_t *in, int shift, uint32_t mask, uint32_t *out,
                                                        int64_t n, const uint32_t *partial) {
    __shared__ uint32_t s_wave[4];
    const int64_t base = (int64_t)blockIdx.x * kChunk + (int64_t)threadIdx.x * kItems;
    uint32_t v[kItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        int64_t i = base + k;
        v[k] = i < n ? scan_in(in, i, shift, mask) : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t pre = block_excl_scan256(s, s_wave, &tot) + partial[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        int64_t i = base + k;
        if (i < n) out[i] = pre;
        pre += v[k];
    }
}

__global__ void __launch_bounds__(kThreads) k_rs_hist(const uint32_t *keys, int64_t n, int shift, uint32_t dmask,
                                                      uint32_t *hist, int64_t ntiles) {
    __shared__ uint32_t h[kRadix];
    const int t = threadIdx.x;
    h[t] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kChunk;
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        int64_t e = base + r * kThreads + t;
        if (e < n) atomicAdd(&h[(keys[e] >> shift) & dmask], 1u);
    }
    __syncthreads();
    hist[(int64_t)t * ntiles + blockIdx.x] = h[t];
}

__global__ void __launch_bounds__(kThreads) k_rs_scatter(const uint32_t *keys_in, const uint32_t *vals_in, int64_t n,
                                                         int shift, uint32_t dmask, const uint32_t *offsets,
                                                         int64_t ntiles,
                                                         uint32_t *keys_out, uint32_t *vals_out, uint32_t *inv_out) {
    __shared__ uint32_t s_keys[kChunk];
    __shared__ uint32_t s_vals[kChunk];
    __shared__ uint32_t s_cnt[kRadix];
    __shared__ uint32_t s_wcnt[4][kRadix];
    __shared__ uint32_t s_start[kRadix];
    __shared__ uint32_t s_wave[4];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t tile = blockIdx.x, base = tile * kChunk;
    s_cnt[t] = 0;
    uint32_t mkey[kItems], mval[kItems], mrank[kItems];
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const int64_t e = base + r * kThreads + t;
        const bool valid = e < n;
        const uint32_t key = valid ? keys_in[e] : 0u;
        const uint32_t val = valid ? (vals_in ? vals_in[e] : (uint32_t)e) : 0u;
        const uint32_t d = (key >> shift) & dmask;
#pragma unroll
        for (int k = 0; k < 4; ++k) s_wcnt[k][t] = 0;
        __syncthreads();
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t rank_w = (uint32_t)__popcll(m & lt_mask);
        if (valid && rank_w == 0) s_wcnt[w][d] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t pre = s_cnt[d];
        for (int k = 0; k < w; ++k) pre += s_wcnt[k][d];
        mkey[r] = key;
        mval[r] = val;
        mrank[r] = valid ? pre + rank_w : 0xffffffffu;
        __syncthreads();
        s_cnt[t] += s_wcnt[0][t] + s_wcnt[1][t] + s_wcnt[2][t] + s_wcnt[3][t];
    }
    __syncthreads();
    uint32_t tot;
    s_start[t] = block_excl_scan256(s_cnt[t], s_wave, &tot);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        if (mrank[r] != 0xffffffffu) {
            const uint32_t d = (mkey[r] >> shift) & dmask;
            const uint32_t slot = s_start[d] + mrank[r];
            s_keys[slot] = mkey[r];
            s_vals[slot] = mval[r];
        }
    }
    __syncthreads();
    const int64_t tile_n = (n - base) < kChunk ? (n - base) : kChunk;
    for (int s = t; s < tile_n; s += kThreads) {
        const uint32_t key = s_keys[s], val = s_vals[s];
        const uint32_t d = (key >> shift) & dmask;
        const uint32_t pos = offsets[(int64_t)d * ntiles + tile] + (uint32_t)(s - s_start[d]);
        keys_out[pos] = key;
        vals_out[pos] = val;
        if (inv_out) inv_out[val] = pos;
    }
}

inline int64_t div_up(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

void fill_u32(uint32_t *p, uint32_t value, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = div_up(n, kThreads);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(kThreads), 0, s, p, value, n);
}

size_t scan_scratch_bytes(int64_t n) { return scan_scratch_bound(n); }

void exclusive_scan_u32(const uint32_t *in, uint32_t *out, int64_t n, void *scratch, hipStream_t s, int shift,
                        uint32_t mask) {
    if (n <= 0) return;
    const int64_t nchunks = div_up(n, kChunk);
    uint32_t *partial = (uint32_t *)scratch;
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nchunks), dim3(kThreads), 0, s, in, n, shift, mask, partial);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, s, partial, nchunks);
    hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nchunks), dim3(kThreads), 0, s, in, shift, mask, out, n,
                       (const uint32_t *)partial);
}

size_t radix_scratch_bytes(int64_t n) { return radix_scratch_bound(n); }

void radix_sort_pairs(const uint32_t *keys_in, const uint32_t *vals_in, int64_t n, int bit_lo, int bit_hi,
                      uint32_t *keys_out, uint32_t *vals_out, uint32_t *inv_out, void *scratch, hipStream_t s) {
    if (n <= 0) return;
    const int64_t ntiles = div_up(n, kChunk);
    char *p = (char *)scratch;
    uint32_t *tk = (uint32_t *)p;
    p += align256(sizeof(uint32_t) * (size_t)n);
    uint32_t *tv = (uint32_t *)p;
    p += align256(sizeof(uint32_t) * (size_t)n);
    uint32_t *hist = (uint32_t *)p;
    p += align256(sizeof(uint32_t) * (size_t)(ntiles * kRadix));
    uint32_t *offs = (uint32_t *)p;
    p += align256(sizeof(uint32_t) * (size_t)(ntiles * kRadix));
    void *scan_scratch = p;
    int passes = (bit_hi - bit_lo + 7) / 8;
    if (passes < 1) passes = 1;
    // the passes alternate between the scratch pair and the outputs, so that the last one
    // writes the outputs: pass k writes the outputs when passes - 1 - k is even
    for (int pass = 0; pass < passes; ++pass) {
        const bool to_out = ((passes - 1 - pass) & 1) == 0;
        const uint32_t *ksrc = pass == 0 ? keys_in : (to_out ? tk : keys_out);
        const uint32_t *vsrc = pass == 0 ? vals_in : (to_out ? tv : vals_out);
        const bool last = pass == passes - 1;
        uint32_t *kdst = to_out ? keys_out : tk;
        uint32_t *vdst = to_out ? vals_out : tv;
        const int shift = bit_lo + 8 * pass;
        const int bits = bit_hi - shift < 8 ? bit_hi - shift : 8;
        const uint32_t dmask = bits <= 0 ? 0u : ((1u << bits) - 1u);
        hipLaunchKernelGGL(k_rs_hist, dim3((unsigned)ntiles), dim3(kThreads), 0, s, ksrc, n, shift, dmask, hist,
                           ntiles);
        exclusive_scan_u32(hist, offs, ntiles * kRadix, scan_scratch, s, 0, 0xffffffffu);
        hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)ntiles), dim3(kThreads), 0, s, ksrc, vsrc, n, shift, dmask,
                           (const uint32_t *)offs, ntiles, kdst, vdst, last ? inv_out : nullptr);
    }
}

}  // namespace zmpm
