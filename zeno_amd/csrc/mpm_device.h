// Device helpers shared by the libzmpm kernel translation units: the resident
// particle tile layout, scene lookups, B-spline weights, the wave-level LDS
// ordering barrier, grid velocity staging and the G2P stencil contraction.
#pragma once

#include "zmpm_internal.h"
#include "mpm_math.h"

namespace zmpm {

constexpr int kThreads = 256;

// Resident particle state: tiles of 64 particles, channel-major inside a tile:
// channel c of particle i at ch[((i / 64) * kTileCh + c) * 64 + i % 64].  A wave's
// 64 lanes read 256 contiguous bytes per channel, and all channels of a particle
// are immediate offsets from one address.
__device__ __forceinline__ uint32_t tile_off(uint32_t i) { return (i >> 6) * (uint32_t)(kTileCh * 64) + (i & 63u); }
__device__ __forceinline__ float &chan(float *ch, uint32_t i, int c) { return ch[tile_off(i) + c * 64]; }
__device__ __forceinline__ const float &chan(const float *ch, uint32_t i, int c) { return ch[tile_off(i) + c * 64]; }
__device__ __forceinline__ uint32_t chan_u(const float *ch, uint32_t i, int c) { return __float_as_uint(chan(ch, i, c)); }
__device__ __forceinline__ void set_u(float *ch, uint32_t i, int c, uint32_t v) { chan(ch, i, c) = __uint_as_float(v); }

// A run of particles starting at a uniform index (a P2G chunk): a 64-bit base at
// the run's first tile, uniform across the wave, and 32-bit byte offsets per lane
// from it, so a channel load takes the scalar-base form (base in SGPRs, lane
// offset in one VGPR, channel as an immediate) instead of a 64-bit address per
// channel and lane.  The run must span less than 4 GiB of tiles.
struct TileRun {
    const float *base;
    uint32_t t0;
};
__device__ __forceinline__ TileRun tile_run(const float *ch, uint32_t first) {
    const uint32_t t0 = first >> 6;
    return {ch + (size_t)t0 * (kTileCh * 64), t0};
}
// Byte offset of particle i (i >= the run's first) from the run's base.
__device__ __forceinline__ uint32_t run_off(const TileRun &r, uint32_t i) {
    return ((i >> 6) - r.t0) * (uint32_t)(kTileCh * 64 * 4) + (i & 63u) * 4u;
}
__device__ __forceinline__ float run_chan(const TileRun &r, uint32_t boff, int c) {
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(r.base) + (size_t)boff + (size_t)c * 256);
}

__device__ __forceinline__ int clamp03(int v) { return v < 0 ? 0 : (v > 3 ? 3 : v); }

// Lattice coordinates (per axis, in blocks from the origin) of lattice index l
// inside its scene's part of the lattice.
__device__ __forceinline__ void lattice_coords(const SceneParams &sc, int64_t l, int &lx, int &ly, int &lz) {
    const int nb = sc.nb;
    const int li = (int)(l - sc.lat_off);
    lz = li % nb;
    ly = (li / nb) % nb;
    lx = li / (nb * nb);
}

// Parameters of scene `scene` (no memory access in a one-scene context).
__device__ __forceinline__ SceneParams scene_params(const Lattice &L, uint32_t scene) {
    return L.nscenes == 1 ? L.s0 : L.scenes[scene];
}

// Scene owning lattice entry l (binary search over the scenes' first entries).
__device__ __forceinline__ int scene_of(const Lattice &L, int64_t l) {
    int lo = 0, hi = L.nscenes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (L.scenes[mid].lat_off <= l) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Quadratic B-spline weights of one axis from the fractional position f
// (the same expressions as make_arena).
__device__ __forceinline__ void bspline_w(float f, float w[3]) {
    w[0] = 0.5f * (1.5f - f) * (1.5f - f);
    w[1] = 0.75f - (f - 1.0f) * (f - 1.0f);
    w[2] = 0.5f * (f - 0.5f) * (f - 0.5f);
}

// Orders this wave's LDS accesses: nothing moves across it, so an LDS store
// issued before it is seen by every lane's LDS load issued after it (a wave's
// LDS instructions execute in issue order).
__device__ __forceinline__ void wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// First logical item (of n) of this workgroup in XCD slab order.  The hardware
// deals workgroup g to XCD g % 8, each XCD with its own L2; with at least
// 8 ceil(n / 8) workgroups, XCD k takes items [k per, (k + 1) per), a slab of
// neighbouring blocks that read the same grid tiles, and the workgroups past
// 8 per take none (n is returned).  With fewer workgroups: the plain order.
// A bijection either way, so a loop "for (b = xcd_first(n); b < n; b += gridDim.x)"
// visits every item once.
__device__ __forceinline__ uint32_t xcd_first(uint32_t n) {
    const uint32_t per = (n + 7u) >> 3;
    if (gridDim.x < 8u * per) return blockIdx.x;
    return blockIdx.x < 8u * per ? (blockIdx.x & 7u) * per + (blockIdx.x >> 3) : n;
}

// The parts of a partition (what a part-indexed loop reads of it); shift: the part size the
// partition was cut with (Lattice::part_shift of the launch).
struct PartView {
    const uint32_t *pblk_start, *pblk_x, *part_blk, *part_lo;
    uint32_t parts_cap, shift;
};
__host__ __device__ __forceinline__ PartView part_view(const Partition &P, uint32_t shift) {
    return PartView{P.pblk_start, P.pblk_x, P.part_blk, P.part_lo, P.parts_cap, shift};
}

// Work unit j (< np + nx) of a partition with np particle blocks and nx extra parts: its block
// b, the index p of its tile and move counts (b for a block's first part), its particles
// [lo, hi) (at most 2^shift), and for an extra part its rank k among all extra parts
// (0xffffffff for a first part).  A first part reads nothing beyond the block starts.
struct PartRef {
    uint32_t b, p, lo, hi, k;
};
__device__ __forceinline__ PartRef part_ref(const PartView &V, uint32_t j, uint32_t np) {
    PartRef r;
    if (j < np) {  // both starts loaded at once
        r.b = r.p = j;
        r.k = 0xffffffffu;
        r.lo = V.pblk_start[j];
        r.hi = min(V.pblk_start[j + 1], r.lo + (1u << V.shift));  // lo < 2^27: no overflow
    } else {
        r.k = j - np;
        r.p = V.parts_cap - 1u - r.k;
        r.b = V.part_blk[r.p];
        r.lo = V.part_lo[r.p];
        r.hi = min(V.pblk_start[r.b + 1], r.lo + (1u << V.shift));
    }
    return r;
}

// Set bits of the wave mask m in the lanes below this one (v_mbcnt: no per-lane
// mask register).
__device__ __forceinline__ uint32_t count_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// The valid lanes of this wave whose key (its low BITS bits) equals this lane's,
// for every lane at once (call it with the whole wave): per bit, the bit's ballot
// where this lane's bit is set and its complement where it is clear, kept as two
// 32-bit halves so that each bit costs one bitwise select per half.
template <int BITS>
__device__ __forceinline__ uint64_t same_key_lanes(uint32_t key, bool valid) {
    const uint64_t vb = __ballot(valid);
    uint32_t lo = (uint32_t)vb, hi = (uint32_t)(vb >> 32);
#pragma unroll
    for (int bit = 0; bit < BITS; ++bit) {
        const uint32_t on = (key >> bit) & 1u;
        const uint64_t bb = __ballot(on != 0u);
        const uint32_t flip = on - 1u;  // 0 where the bit is set, all ones where it is clear
        lo &= (uint32_t)bb ^ flip;
        hi &= (uint32_t)(bb >> 32) ^ flip;
    }
    return ((uint64_t)hi << 32) | lo;
}

// HBM layout of a particle block's P2G tile (channel-major, kTileNodes floats per
// channel): its 6^3 nodes [K, K+6)^3 as the eight regions that fall into the grid
// blocks B + e, e = (ex, ey, ez) in {0,1}^3, one after another (region e holds
// (4 - 2 ex) x (4 - 2 ey) x (4 - 2 ez) nodes, x-major), so the grid update reads
// the part of each tile that covers its block as one contiguous run.
__device__ __forceinline__ int tile_region_offset(int e) {
    // region sizes by e = 4 ex + 2 ey + ez: 64, 32, 32, 16, 32, 16, 16, 8
    return e == 0 ? 0 : e == 1 ? 64 : e == 2 ? 96 : e == 3 ? 128 : e == 4 ? 144 : e == 5 ? 176 : e == 6 ? 192 : 208;
}
// Slot of tile node (x, y, z), each in [0, 6).
__device__ __forceinline__ int tile_slot(int x, int y, int z) {
    const int ex = x >> 2, ey = y >> 2, ez = z >> 2;
    const int sy = 4 - 2 * ey, sz = 4 - 2 * ez;
    return tile_region_offset(4 * ex + 2 * ey + ez) + ((x - 4 * ex) * sy + (y - 4 * ey)) * sz + (z - 4 * ez);
}
// Slot of LDS tile node k = (x * 6 + y) * 6 + z.
__device__ __forceinline__ int tile_slot_of(int k) { return tile_slot(k / 36, (k / 6) % 6, k % 6); }

// Grid velocities of one particle block's 6^3 stencil region, one float4 (v, 0)
// per node so a z-run of 3 nodes is 3 contiguous 16-byte LDS reads.
template <int NT = kThreads>
__device__ __forceinline__ void stage_block_velocity(float4 *s_v, const float *grid, const int32_t *lat_gid, int64_t l,
                                                     int64_t nb, int64_t nb2) {
    for (int k = threadIdx.x; k < kTileNodes; k += NT) {
        const int tx = k / 36, ty = (k / 6) % 6, tz = k % 6;
        const int ex = tx >> 2, ey = ty >> 2, ez = tz >> 2;
        const int32_t g = lat_gid[l + ex * nb2 + ey * nb + ez];
        const int cell = ((tx & 3) * 4 + (ty & 3)) * 4 + (tz & 3);
        const float *gv = grid + (int64_t)(g >= 0 ? g : 0) * 256 + 64 + cell;
        const float v0 = gv[0], v1 = gv[64], v2 = gv[128];
        s_v[k] = g >= 0 ? make_float4(v0, v1, v2, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// v_p = sum W v_i and C_p = D^-1 sum W v_i d_i^T over the 27 stencil nodes
// (Transfer.cu:1145-1162), contracted one axis at a time: with W = wx wy wz and
// d_i = ((ox, oy, oz) - f) dx, the z-sums a = sum wz v and az = sum wz dz v over
// each z-run, then y-sums of those, then x-sums, give v and the three columns of
// C with 297 multiply-adds instead of 486.  v is formed by the same sequence with
// or without C, so both G2P passes agree bit for bit on the advected position.
// grad (non-affine FLIP, Transfer.cu:1232-1240): the weight derivative dw/df
// (f - 3/2, 2 (1 - f), f - 1/2) in place of w d, so Cn = Dinv * sum v_i (dW/df)^T
// is the velocity gradient sum v_i grad W^T when Dinv is 1/dx.
// get(ox, oy, oz): the velocity (float4, w unused) of stencil node (ox, oy, oz)
// from the LDS-staged block (g2p_gather) or from the grid in HBM (arrivals of
// the fused pass); the arithmetic is the same either way.
// ROLL: one x-offset at a time (the 27 node loads of the unrolled contraction are
// issued first and hold 108 registers; the sums are the same either way).  Only
// where registers bound the kernel: rolled, the APIC update pass was 40 % slower
// and the AFLIP one (252 registers unrolled) 6 % slower.
template <bool WITH_C, bool ROLL = false, class Get>
__device__ __forceinline__ void g2p_gather_fn(Get get, const Arena &ar, float dx, float Dinv, float vel[3],
                                              float Cn[3][3], bool grad = false) {
    float wd[3][3];  // w * d (or dw/df) per axis and offset
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float dw[3] = {ar.f[k] - 1.5f, 2.0f * (1.0f - ar.f[k]), ar.f[k] - 0.5f};
#pragma unroll
        for (int o = 0; o < 3; ++o) wd[k][o] = grad ? dw[o] : ar.w[k][o] * (((float)o - ar.f[k]) * dx);
    }
    float cx[3], cy[3], cz[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) vel[r] = cx[r] = cy[r] = cz[r] = 0.0f;
    constexpr int kUnrollX = ROLL ? 1 : 3;
#pragma unroll kUnrollX
    for (int ox = 0; ox < 3; ++ox) {
        float b[3] = {0.0f, 0.0f, 0.0f}, by[3] = {0.0f, 0.0f, 0.0f}, bz[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int oy = 0; oy < 3; ++oy) {
            float za[3] = {0.0f, 0.0f, 0.0f}, zd[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int oz = 0; oz < 3; ++oz) {
                const float4 v4 = get(ox, oy, oz);
                const float vi[3] = {v4.x, v4.y, v4.z};
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    za[r] = fmaf(ar.w[2][oz], vi[r], za[r]);
                    if (WITH_C) zd[r] = fmaf(wd[2][oz], vi[r], zd[r]);
                }
            }
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                b[r] = fmaf(ar.w[1][oy], za[r], b[r]);
                if (WITH_C) {
                    by[r] = fmaf(wd[1][oy], za[r], by[r]);
                    bz[r] = fmaf(ar.w[1][oy], zd[r], bz[r]);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            vel[r] = fmaf(ar.w[0][ox], b[r], vel[r]);
            if (WITH_C) {
                cx[r] = fmaf(wd[0][ox], b[r], cx[r]);
                cy[r] = fmaf(ar.w[0][ox], by[r], cy[r]);
                cz[r] = fmaf(ar.w[0][ox], bz[r], cz[r]);
            }
        }
    }
    if (WITH_C) {
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            Cn[r][0] = Dinv * cx[r];
            Cn[r][1] = Dinv * cy[r];
            Cn[r][2] = Dinv * cz[r];
        }
    }
}

// g2p_gather_fn on the 6^3 nodes staged in LDS; lc: the particle's cell in its block.
template <bool WITH_C, bool ROLL = false>
__device__ __forceinline__ void g2p_gather(const float4 *s_v, const Arena &ar, const int lc[3], float dx, float Dinv,
                                           float vel[3], float Cn[3][3], bool grad = false) {
    const float4 *base = s_v + (lc[0] * kTileSide + lc[1]) * kTileSide + lc[2];
    g2p_gather_fn<WITH_C, ROLL>([&](int ox, int oy, int oz) { return base[(ox * kTileSide + oy) * kTileSide + oz]; }, ar,
                          dx, Dinv, vel, Cn, grad);
}

// Velocity of grid node (cx, cy, cz) (cell coordinates) straight from the grid in
// HBM: its block's id through the lattice map, then the node's channels.  Every
// stencil node of a particle lies in one of its block's grid blocks B + {0,1}^3,
// which exist; the clamps only keep a bad position inside the buffers.
__device__ __forceinline__ float4 grid_node_velocity(const float *grid, const int32_t *lat_gid, const SceneParams &sc,
                                                     int origin, int cx, int cy, int cz) {
    const int nb = sc.nb;
    int bx = (cx >> 2) - origin, by = (cy >> 2) - origin, bz = (cz >> 2) - origin;
    bx = bx < 0 ? 0 : (bx >= nb ? nb - 1 : bx);
    by = by < 0 ? 0 : (by >= nb ? nb - 1 : by);
    bz = bz < 0 ? 0 : (bz >= nb ? nb - 1 : bz);
    const int32_t g = lat_gid[sc.lat_off + (bx * nb + by) * nb + bz];
    const int cell = ((cx & 3) * 4 + (cy & 3)) * 4 + (cz & 3);
    const float *gv = grid + (int64_t)(g >= 0 ? g : 0) * 256 + 64 + cell;
    const float v0 = gv[0], v1 = gv[64], v2 = gv[128];
    return g >= 0 ? make_float4(v0, v1, v2, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

}  // namespace zmpm
