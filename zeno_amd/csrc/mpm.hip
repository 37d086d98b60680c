// MLS/APIC-MPM substep kernels for gfx950 (wave64, 160 KB LDS per CU).
//
// One workgroup per 4^3 particle block (one wave per grid block in the grid
// kernel); every launch covers all scenes of a context, each block reading its
// scene's parameters.  No float atomics anywhere: every result is deterministic.
//   bin keys      one lane per particle: key = (lattice index << 6) | cell id of
//                 floor(x * (1/dx) - 0.5) (MPMPipeline.cu:75-84).
//   P2G           per particle block (3 waves): cells and stable in-cell ranks by
//                 wave ballots, stress (FP64 polar for fixed corotated) and a
//                 16-float affine payload staged in LDS in cell order; lane = cell,
//                 wave = stencil x-offset accumulation in registers; private 6^3
//                 LDS tiles summed in a fixed order into the block's HBM tile.
//   grid          one wave per grid block: the <= 8 overlapping tiles summed in a
//                 fixed order, v = mv/m + a dt, sticky box, per-block max |v|^2.
//   G2P move      per particle block: velocities staged in LDS, each particle's
//                 direction (1 of 27) to the block of x + v dt and its stable rank.
//   lattice scan  decoupled look-back over the occupied lattice tiles: the next
//                 partition (blocks in lattice order, first particle of each).
//   G2P update    per particle block: v, C, x, F|J, written straight into the
//                 next stable block order.
// Reference call sites: P2G Transfer.cu:844-890, grid update MPMPipeline.cu:852-877,
// boundary Boundary.cu:104-116, G2P Transfer.cu:1137-1173.
#include "zmpm_internal.h"
#include "mpm_math.h"
#include "mpm_device.h"
#include "../../include/zmpm.h"

namespace zmpm {

namespace {

// Key of a position in scene sc: (lattice index << 6) | cell id.  Positions outside
// the scene's lattice (or not finite) are clamped onto it and reported through `bad`.
__device__ __forceinline__ uint32_t position_key(const float xp[3], const SceneParams &sc, int origin, int key_shift,
                                                 bool &bad) {
    const float dxinv = sc.dxinv;
    const int nbs = sc.nb;
    int lb[3], cl[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float X = xp[k] * dxinv;
        const double c = (double)X - 0.5;
        int cell;
        if (c > -1.0e9 && c < 1.0e9) {
            cell = (int)floor(c);
        } else {
            cell = origin * 4;
            bad = true;
        }
        int b = (cell >> 2) - origin;  // floor(cell / 4) for two's complement cells
        if (b < 0 || b > nbs - 2) {
            bad = true;
            b = b < 0 ? 0 : nbs - 2;
        }
        lb[k] = b;
        cl[k] = cell & 3;
    }
    const int64_t li = (int64_t)sc.lat_off + ((int64_t)lb[0] * nbs + lb[1]) * nbs + lb[2];
    // lattices of 2^26 entries or more: the lattice index alone (Lattice::key_shift 0)
    return key_shift ? (uint32_t)(li << 6) | (uint32_t)((cl[0] * 4 + cl[1]) * 4 + cl[2]) : (uint32_t)li;
}

// ------------------------------------------------------------------ keys
// x: the resident tiled state (stride < 0; scene from the material) or SoA
// planes of `stride` entries (scene 0).
__device__ __forceinline__ uint32_t particle_key(const float *x, int64_t stride, int64_t i, const DevMaterial *mats,
                                                 const Lattice &L, bool &bad) {
    float xp[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) xp[k] = stride < 0 ? chan(x, (uint32_t)i, kChX + k) : x[k * stride + i];
    const int scene = stride < 0 ? mats[chan_u(x, (uint32_t)i, kChMat) & kMatIndexMask].scene : 0;
    return position_key(xp, scene_params(L, (uint32_t)scene), L.origin, L.key_shift, bad);
}

// check: also flag (bin_mismatch) a state whose blocks are out of order, which a
// state assumed to be in stable block order must not be.
__global__ void __launch_bounds__(kThreads) k_bin_keys(const float *x, int64_t stride, int64_t n,
                                                       const DevMaterial *mats, Lattice L, uint32_t *keys,
                                                       DevStatus *st, int check) {
    bool bad = false, unsorted = false;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        const uint32_t key = particle_key(x, stride, i, mats, L, bad);
        keys[i] = key;
        if (check && i > 0)
            unsorted |= (particle_key(x, stride, i - 1, mats, L, bad) >> L.key_shift) > (key >> L.key_shift);
    }
    if (bad) atomicOr(&st->domain_error, 1u);
    if (unsorted) atomicOr(&st->bin_mismatch, 1u);
}

// Resets the lattice scan's look-back state; force: make the scan visit every tile.
__global__ void __launch_bounds__(kThreads) k_scan_reset(Binning B, uint32_t force) {
    for (int64_t k = blockIdx.x * (int64_t)kThreads + threadIdx.x; k < B.scan_tiles;
         k += (int64_t)gridDim.x * kThreads) {
        B.scan_p.status[k] = 0ull;
        B.scan_g.status[k] = 0ull;
        B.tile_mark[k] = force;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *B.scan_p.tile_ctr = 0u;
}

// ------------------------------------------------------------------ partition
// Full re-bin path: mark the particle blocks of a key array sorted by block and
// the grid blocks B + {0,1}^3 their stencils reach (ExpandZSPartition,
// MPMPipeline.cu:131-199, restricted to the directions a quadratic stencil
// touches).  Lattice coordinates of a particle block are <= nb-2 by
// construction, so the +1 neighbours exist.
// Each block's particle count goes to dcount (one atomic per run of one block within a
// wave), from which the lattice scan forms the block starts and parts, as it does for the
// counts of the move pass.
__global__ void __launch_bounds__(kThreads) k_mark_blocks(const uint32_t *skeys, int64_t n, Lattice L,
                                                          const float *ch, const DevMaterial *mats,
                                                          uint32_t *lat_flag, uint32_t *dcount) {
    const int lane = threadIdx.x & 63;
    for (int64_t k0 = blockIdx.x * (int64_t)kThreads; k0 < n; k0 += (int64_t)gridDim.x * kThreads) {  // uniform
        const int64_t k = k0 + threadIdx.x;
        const bool valid = k < n;
        const uint32_t blk = valid ? skeys[k] >> L.key_shift : 0xffffffffu;
        const uint32_t prev = __shfl_up(blk, 1, 64);
        if (valid && (k == 0 || (skeys[k - 1] >> L.key_shift) != blk)) {
            const int64_t nb = L.scenes[mats[chan_u(ch, (uint32_t)k, kChMat) & kMatIndexMask].scene].nb;
            const int64_t nb2 = nb * nb;
#pragma unroll
            for (int e = 0; e < 8; ++e)
                atomicOr(&lat_flag[blk + (e >> 2) * nb2 + ((e >> 1) & 1) * nb + (e & 1)], e == 0 ? 3u : 2u);
        }
        // a run of this wave's lanes in one block: from its first lane to the next run's first
        // lane or past the wave's last valid lane (valid lanes are a prefix of the wave)
        const bool run_start = valid && (lane == 0 || prev != blk);
        const uint64_t starts = __ballot(run_start);
        const uint64_t live = __ballot(valid);
        if (run_start) {
            const uint64_t later = lane == 63 ? 0ull : (starts >> (lane + 1)) << (lane + 1);
            const int end = later ? __ffsll((long long)later) - 1 : 64 - __clzll((long long)live);
            atomicAdd(&dcount[blk], (uint32_t)(end - lane));
        }
    }
}

// Lattice scan: two decoupled look-back chains over the lattice in index
// order give every marked entry its particle-block id (and the first
// particle of that block, from the per-block particle counts) and its
// grid-block id.  Clears the marks and counts it read.  A tile of 4096 entries
// is read and written only if it may hold a new mark or holds an entry of this
// partition buffer's previous contents (Pn.tile_used); otherwise it contributes
// zero and its entries are already -1, so a sparse lattice costs its occupied
// tiles, not its size.  New marks lie within B.mark_radius tiles of a tile
// holding a block of the partition the particles move from (prev_used; a
// block marks entries at offsets -1 .. +2 per axis), or anywhere when
// B.tile_mark forces a tile (full re-bin).  Entries move as 16-byte vectors.
// The grid-block chain also counts extra parts (bits 32-61: a particle block of d
// particles has ceil(d / 2^L.part_shift) - 1 of them), so every block gets the number of
// extra parts before it and every extra part its block and first particle.
__device__ __forceinline__ uint32_t extra_parts(uint32_t d, uint32_t shift) { return d == 0u ? 0u : (d - 1u) >> shift; }

__global__ void __launch_bounds__(kScanThreads) k_lattice_scan(Lattice L, Partition Pn, Binning B,
                                                              const uint32_t *prev_used, DevStatus *st) {
    __shared__ uint64_t s_wave[4];
    __shared__ uint32_t s_tile;
    // particle-block count in the low kPBlkBits bits (blocks <= capacity < 2^27), particles above
    constexpr uint64_t kBlkMask = (1ull << kPBlkBits) - 1;
    if (threadIdx.x == 0) s_tile = atomicAdd(B.scan_p.tile_ctr, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const int64_t size = L.size;
    uint32_t near = B.tile_mark[tile] | Pn.tile_used[tile];
    {
        const int64_t lo = (int64_t)tile - B.mark_radius, hi = (int64_t)tile + B.mark_radius;
        for (int64_t q = lo < 0 ? 0 : lo; q <= hi && q < B.scan_tiles; ++q) near |= prev_used[q];
    }
    const bool need = near != 0u;
    const int64_t base = (int64_t)tile * kScanTile + (int64_t)threadIdx.x * kScanItems;
    const bool full = base + kScanItems <= size;
    uint32_t f[kScanItems], d[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) f[k] = d[k] = 0u;
    if (need) {
        if (full) {
#pragma unroll
            for (int q = 0; q < kScanItems / 4; ++q) {
                const uint4 fv = *reinterpret_cast<const uint4 *>(B.lat_flag + base + 4 * q);
                const uint4 dv = *reinterpret_cast<const uint4 *>(B.dcount + base + 4 * q);
                f[4 * q] = fv.x, f[4 * q + 1] = fv.y, f[4 * q + 2] = fv.z, f[4 * q + 3] = fv.w;
                d[4 * q] = dv.x, d[4 * q + 1] = dv.y, d[4 * q + 2] = dv.z, d[4 * q + 3] = dv.w;
            }
        } else {
#pragma unroll
            for (int k = 0; k < kScanItems; ++k)
                if (base + k < size) {
                    f[k] = B.lat_flag[base + k];
                    d[k] = B.dcount[base + k];
                }
        }
    }
    uint64_t sp = 0;
    uint32_t sgb = 0, sgx = 0;  // grid blocks, extra parts
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        sp += ((uint64_t)d[k] << kPBlkBits) | (uint64_t)(f[k] & 1u);
        sgb += (f[k] >> 1) & 1u;
        sgx += extra_parts(d[k], L.part_shift);  // 0 for entries without particles
    }
    const uint64_t sg = ((uint64_t)sgx << 32) | sgb;
    uint64_t totp = 0, totg = 0, ep = 0, eg = 0;
    if (need) {
        ep = block_excl_scan_u64(sp, s_wave, &totp);
        eg = block_excl_scan_u64(sg, s_wave, &totg);
    }
    // the two chains look back concurrently, one wave each
    __shared__ uint64_t s_pref[2];
    if (threadIdx.x < 64) {
        const uint64_t x = lookback_wave(B.scan_p, tile, totp);
        if (threadIdx.x == 0) s_pref[0] = x;
    } else if (threadIdx.x < 128) {
        const uint64_t x = lookback_wave(B.scan_g, tile, totg);
        if (threadIdx.x == 64) s_pref[1] = x;
    }
    __syncthreads();
    const uint64_t tp = s_pref[0], tg = s_pref[1];
    if (need) {
        uint64_t rp = tp + ep;
        uint32_t rg = (uint32_t)(tg + eg), rx = (uint32_t)((tg + eg) >> 32);
        // scene of this thread's first entry; later entries advance from it
        int scene = base < size ? scene_of(L, base) : 0;
        int32_t op[kScanItems], og[kScanItems];
        uint32_t xf[kScanItems / 4] = {0u, 0u, 0u, 0u};  // extra-part flags, one byte per entry
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const int64_t i = base + k;
            op[k] = og[k] = -1;
            if (i >= size) continue;
            while (scene + 1 < L.nscenes && L.scenes[scene + 1].lat_off <= i) ++scene;
            const uint32_t extra = extra_parts(d[k], L.part_shift);
            if (f[k] & 1u) {
                const uint32_t pidx = (uint32_t)(rp & kBlkMask), first = (uint32_t)(rp >> kPBlkBits);
                op[k] = (int32_t)pidx;
                Pn.pblk_lat[pidx] = (uint32_t)i;
                Pn.pblk_start[pidx] = first;
                Pn.pblk_scene[pidx] = (uint32_t)scene;
                Pn.pblk_x[pidx] = rx;
                if (extra) xf[k / 4] |= 1u << (8 * (k % 4));
                for (uint32_t q = 1; q <= extra; ++q) {
                    const uint32_t e = Pn.parts_cap - 1u - (rx + q - 1u);
                    Pn.part_blk[e] = pidx;
                    Pn.part_lo[e] = first + (q << L.part_shift);
                }
            }
            if ((f[k] >> 1) & 1u) {
                og[k] = (int32_t)rg;
                Pn.gblk_lat[rg] = (uint32_t)i;
                Pn.gblk_scene[rg] = (uint32_t)scene;
            }
            rp += ((uint64_t)d[k] << kPBlkBits) | (uint64_t)(f[k] & 1u);
            rg += (f[k] >> 1) & 1u;
            rx += extra;
        }
        if (full) {
#pragma unroll
            for (int q = 0; q < kScanItems / 4; ++q) {
                *reinterpret_cast<int4 *>(Pn.lat_pid + base + 4 * q) =
                    make_int4(op[4 * q], op[4 * q + 1], op[4 * q + 2], op[4 * q + 3]);
                *reinterpret_cast<int4 *>(Pn.lat_gid + base + 4 * q) =
                    make_int4(og[4 * q], og[4 * q + 1], og[4 * q + 2], og[4 * q + 3]);
                if (q == 0) *reinterpret_cast<uint4 *>(Pn.lat_xf + base) = make_uint4(xf[0], xf[1], xf[2], xf[3]);
                if (f[4 * q] | f[4 * q + 1] | f[4 * q + 2] | f[4 * q + 3]) {
                    *reinterpret_cast<uint4 *>(B.lat_flag + base + 4 * q) = make_uint4(0u, 0u, 0u, 0u);
                    *reinterpret_cast<uint4 *>(B.dcount + base + 4 * q) = make_uint4(0u, 0u, 0u, 0u);
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < kScanItems; ++k)
                if (base + k < size) {
                    Pn.lat_pid[base + k] = op[k];
                    Pn.lat_gid[base + k] = og[k];
                    Pn.lat_xf[base + k] = (uint8_t)((xf[k / 4] >> (8 * (k % 4))) & 1u);
                    B.lat_flag[base + k] = 0u;
                    B.dcount[base + k] = 0u;
                }
        }
        if (threadIdx.x == 0) {
            Pn.tile_used[tile] = (totp | totg) != 0u ? 1u : 0u;
            B.tile_mark[tile] = 0u;
        }
    }
    if (tile == (uint32_t)B.scan_tiles - 1 && threadIdx.x == 0) {
        const uint64_t ap = tp + totp, ag = tg + totg;
        Pn.counts[0] = (uint32_t)(ap & kBlkMask);
        Pn.counts[1] = (uint32_t)ag;
        Pn.counts[2] = (uint32_t)(ag >> 32);
        Pn.pblk_start[ap & kBlkMask] = (uint32_t)(ap >> kPBlkBits);
        Pn.pblk_x[ap & kBlkMask] = (uint32_t)(ag >> 32);
        if (B.host_counts) {
            B.host_counts[0] = (uint32_t)(ap & kBlkMask);
            B.host_counts[1] = (uint32_t)ag;
            B.host_counts[2] = (uint32_t)(ag >> 32);
        }
    }
    if (threadIdx.x == 0 && *B.scan_p.stall_flag) atomicOr(&st->scan_stall, 1u);
}

__global__ void __launch_bounds__(kThreads) k_gather_state(State src, int64_t src_cap, State dst, int64_t dst_cap,
                                                           const uint32_t *perm, int64_t n) {
    for (int64_t k = blockIdx.x * (int64_t)kThreads + threadIdx.x; k < n; k += (int64_t)gridDim.x * kThreads) {
        const uint32_t i = perm[k];
#pragma unroll
        for (int c = 0; c < kTileCh; ++c) chan(dst.ch, (uint32_t)k, c) = chan(src.ch, i, c);
    }
}

// ------------------------------------------------------------------ P2G
constexpr int kP2GThreads = 192;             // 3 waves: wave g handles the stencil x-offset ox = g in phase B
constexpr int kP2GRounds = 2;                // phase A rounds of 192 particles
// 384 particles per chunk: 26 KB of LDS, five workgroups per CU (a full block at 8
// particles per cell takes two chunks; each lane still walks its cell in block order)
constexpr int kP2GChunk = kP2GThreads * kP2GRounds;
constexpr int kPL = 16;                      // payload floats: m, f[3], a'[3], B dx[3][3]

// Phase B for x-offset ox: accumulate the cell's staged particles into the 9
// nodes (ox, oy, oz), acc[4 * (3 oy + oz) + channel].  The mass at offset o is
// W m, the momentum W (a' + (B dx) o), with B = m C - dt D^-1 vol P F^T and
// a' = m v - B f dx; (B dx) o is built up as a' + ox (B dx)_x, then by repeated
// addition along oy, oz.
// sorted (optional): the payload of cell-order position s sits at sorted[s].
// The weight of x-offset ox alone, bit for bit bspline_w's: w = a + (b (f - c)) (f - c)
// with (a, b, c) = (0, 1/2, 3/2), (3/4, -1, 1), (0, 1/2, 1/2); the negations are exact, so
// (1/2 (f - 3/2)) (f - 3/2) = (1/2 (3/2 - f)) (3/2 - f) and 3/4 + (-(f - 1)) (f - 1) = 3/4 - (f - 1)^2.
struct XWeight {
    float a, b, c;
};
__device__ __forceinline__ XWeight xweight_of(int ox) {
    return ox == 0 ? XWeight{0.0f, 0.5f, 1.5f} : (ox == 1 ? XWeight{0.75f, -1.0f, 1.0f} : XWeight{0.0f, 0.5f, 0.5f});
}
__device__ __forceinline__ float xweight(const XWeight &k, float f) {
    const float d = f - k.c;
    return k.a + (k.b * d) * d;
}

template <int STRIDE = kP2GChunk>
__device__ __forceinline__ void p2g_accumulate(float (&acc)[36], const float *s_pl, int sb, int se, int ox,
                                               const uint16_t *sorted = nullptr) {
    const float fox = (float)ox;
    const XWeight xk = xweight_of(ox);
    for (int s0 = sb; s0 < se; ++s0) {
        const int s = sorted ? (int)sorted[s0] : s0;
        const float m = s_pl[s];
        float f[3], ap[3], bd[9];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            f[q] = s_pl[(1 + q) * STRIDE + s];
            ap[q] = s_pl[(4 + q) * STRIDE + s];
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) bd[q] = s_pl[(7 + q) * STRIDE + s];
        float wy[3], wz[3];
        bspline_w(f[1], wy);
        bspline_w(f[2], wz);
        const float wxo = xweight(xk, f[0]);
        float ty[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) ty[r] = fmaf(fox, bd[3 * r], ap[r]);
#pragma unroll
        for (int oy = 0; oy < 3; ++oy) {
            const float wxy = wxo * wy[oy];
            float t[3] = {ty[0], ty[1], ty[2]};
#pragma unroll
            for (int oz = 0; oz < 3; ++oz) {
                const float W = wxy * wz[oz];
                const int j = 3 * oy + oz;
                acc[4 * j] = fmaf(W, m, acc[4 * j]);
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    acc[4 * j + 1 + r] = fmaf(W, t[r], acc[4 * j + 1 + r]);
                    t[r] += bd[3 * r + 2];
                }
            }
#pragma unroll
            for (int r = 0; r < 3; ++r) ty[r] += bd[3 * r + 1];
        }
    }
}

// Phase C for x-offset ox: add the registers into the x-offset's 6^3 tile (one
// float4 (m, mv) per node).  For one (oy, oz) the lanes (distinct cells, one
// x-offset) hit distinct nodes, but cell c at one (oy, oz) can hit the node of
// cell c' at another, so consecutive (oy, oz) are ordered: ORDER() is
// wave_lds_order() when one wave holds all cells of the tile, a workgroup
// barrier when several waves do (plain read-modify-write: LDS float atomics
// are several times slower here).  Every node thus receives its 9 cells in
// (oy, oz) order, whichever lanes hold them.
template <bool WG_ORDER>
__device__ __forceinline__ void p2g_tile_add(const float (&acc)[36], float4 *tile, int lx, int ly, int lz, int ox) {
#pragma unroll
    for (int oy = 0; oy < 3; ++oy)
#pragma unroll
        for (int oz = 0; oz < 3; ++oz) {
            const int node = ((lx + ox) * kTileSide + (ly + oy)) * kTileSide + (lz + oz);
            const int j = 3 * oy + oz;
            float4 v = tile[node];
            v.x += acc[4 * j];
            v.y += acc[4 * j + 1];
            v.z += acc[4 * j + 2];
            v.w += acc[4 * j + 3];
            tile[node] = v;
            if (WG_ORDER) __syncthreads();
            else wave_lds_order();
        }
}

// Stress term of one F-based particle, out = -dt D^-1 vol P(F) F^T
// (Transfer.cu:865-869) with s = -dt D^-1 vol: fixed corotated with the FP64
// polar rotation, (stable) Neo-Hookean in closed form, StVK-Hencky as its
// Kirchhoff stress P F^T from the FP64 left stretch, plus the ARAP fiber term
// (Transfer.cu:866-867) when the material has one (fa: the particle's fiber).
__device__ __forceinline__ void f_stress_term(const DevMaterial &mt, float s, float mu, float lam, const float F[3][3],
                                              const float fa[3], float out[3][3]) {
    if (mt.model == ZMPM_MODEL_STVK_HENCKY) {
        float tau[3][3];
        stvk_hencky_tau(F, mu, lam, tau);
        if (mt.aniso_mu != 0.0f) {  // tau += P_a F^T
            float Pa[3][3] = {};
            arap_fiber_add(F, fa, mt.aniso_mu, Pa);
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) tau[r][c] += Pa[r][0] * F[c][0] + Pa[r][1] * F[c][1] + Pa[r][2] * F[c][2];
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) out[q / 3][q % 3] = s * tau[q / 3][q % 3];
        return;
    }
    float Pk[3][3];
    if (mt.model == ZMPM_MODEL_NEO_HOOKEAN) neo_hookean_piola(F, mu, lam, Pk);
    else if (mt.model == ZMPM_MODEL_STABLE_NEO_HOOKEAN) stable_neo_hookean_piola(F, mu, lam, Pk);
    else fcr_piola(F, mu, lam, Pk);
    if (mt.aniso_mu != 0.0f) arap_fiber_add(F, fa, mt.aniso_mu, Pk);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            out[r][c] = (s * Pk[r][0]) * F[c][0] + (s * Pk[r][1]) * F[c][1] + (s * Pk[r][2]) * F[c][2];
}

// The projection of one plastic particle: F (and the Cam-clay logJp) replaced
// when the trial state lies outside the yield surface.
__device__ __forceinline__ void plastic_project(const DevMaterial &mt, float F[3][3], float &logJp) {
    float Fp[3][3];
#pragma unroll
    for (int q = 0; q < 9; ++q) Fp[q / 3][q % 3] = F[q / 3][q % 3];
    bool changed;
    if (mt.plasticity == ZMPM_PLASTIC_CAM_CLAY) {
        float lj = logJp;
        changed = project_cam_clay(Fp, mt.mu, mt.nacc_M, mt.nacc_bulk, mt.nacc_beta, mt.nacc_xi, lj);
        if (changed) logJp = lj;
    } else {
        changed = project_hencky_strain(Fp, mt.mu, mt.lam, mt.plasticity, mt.dp_alpha, mt.yield_stress);
    }
    if (changed) {
#pragma unroll
        for (int q = 0; q < 9; ++q) F[q / 3][q % 3] = Fp[q / 3][q % 3];
    }
}

// Stress term of the F-based particles, one lane per particle (f_stress_term).
// Computed apart from P2G so that these registers and run times do not hold
// back the transfer's workgroups; J-fluid particles are skipped (P2G forms
// their tau = K J (J - 1) I itself).
// GENERAL: every F-based model and the ARAP fiber term; otherwise the context's
// F-based particles are all fixed corotated without a fiber term, and the code
// (and registers) of the others is left out.
template <bool GENERAL>
__global__ void __launch_bounds__(kThreads) k_stress(State A, int64_t n, Lattice L, const DevMaterial *mats,
                                                     float *sig, int64_t stride) {
    for (int64_t k = blockIdx.x * (int64_t)kThreads + threadIdx.x; k < n; k += (int64_t)gridDim.x * kThreads) {
        const uint32_t i = (uint32_t)k;
        const uint32_t mw = chan_u(A.ch, i, kChMat);
        if (!(mw & kMatF)) continue;
        const DevMaterial mt = mats[mw & kMatIndexMask];
        const SceneParams sc = scene_params(L, (uint32_t)mt.scene);
        const float s = (-sc.dt * sc.Dinv) * (L.pp_mass ? chan(A.ch, i, kChVol) : mt.vol);
        // the particle's own Lame parameters when it carries them (Transfer.cu:861-864)
        const float mu = L.pp_lame ? chan(A.ch, i, kChMu) : mt.mu;
        const float lam = L.pp_lame ? chan(A.ch, i, kChLam) : mt.lam;
        float F[3][3], out[3][3];
#pragma unroll
        for (int q = 0; q < 9; ++q) F[q / 3][q % 3] = chan(A.ch, i, kChF + q);
        if (GENERAL) {
            float fa[3] = {0.0f, 0.0f, 0.0f};
            if (mt.aniso_mu != 0.0f) {
#pragma unroll
                for (int c = 0; c < 3; ++c) fa[c] = chan(A.ch, i, kChFiber + c);
            }
            f_stress_term(mt, s, mu, lam, F, fa, out);
        } else {  // f_stress_term's fixed-corotated branch, operation for operation
            float Pk[3][3];
            fcr_piola(F, mu, lam, Pk);
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    out[r][c] = (s * Pk[r][0]) * F[c][0] + (s * Pk[r][1]) * F[c][1] + (s * Pk[r][2]) * F[c][2];
        }
#pragma unroll
        for (int q9 = 0; q9 < 9; ++q9) sig[q9 * stride + k] = out[q9 / 3][q9 % 3];
    }
}

// The particle channels P2G reads besides x: v, C, the J-fluid's J and the
// material word.  pp: the state pointer when the particles carry their own mass
// and volume (Transfer.cu:857-858), read where the payload uses them (held from
// the start, the two extra registers spill the transfer kernel); else null and
// the material's are used.
struct P2GIn {
    float v[3], C[9], J;
    uint32_t mw;
    const float *pp;
};

__device__ __forceinline__ void p2g_load(const State &A, uint32_t i, P2GIn &in, bool pp_mass) {
#pragma unroll
    for (int k = 0; k < 3; ++k) in.v[k] = chan(A.ch, i, 3 + k);
#pragma unroll
    for (int q = 0; q < 9; ++q) in.C[q] = chan(A.ch, i, 6 + q);
    in.J = chan(A.ch, i, 15);
    in.mw = chan_u(A.ch, i, kChMat);
    in.pp = pp_mass ? A.ch : nullptr;
}

// The same fields through a run of the state (boff: run_off of particle i).
__device__ __forceinline__ void p2g_load(const State &A, const TileRun &run, uint32_t boff, P2GIn &in, bool pp_mass) {
#pragma unroll
    for (int k = 0; k < 3; ++k) in.v[k] = run_chan(run, boff, 3 + k);
#pragma unroll
    for (int q = 0; q < 9; ++q) in.C[q] = run_chan(run, boff, 6 + q);
    in.J = run_chan(run, boff, 15);
    in.mw = __float_as_uint(run_chan(run, boff, kChMat));
    in.pp = pp_mass ? A.ch : nullptr;
}

// The 16-float P2G payload of particle i, stored at pl[k * STRIDE] (k = 0..15):
// m, f (fractional position in cells), a' = m v - B f dx and B dx, with
// B = m C + the stress term (k_stress's for fixed corotated, tau = K J (J - 1) I
// times -dt D^-1 vol for J-fluid) (Transfer.cu:855-887).
// PART (the FLIP transfers scatter momentum and stress to separate grid
// channels, Transfer.cu:927-930, 977-981): 0 both, 1 momentum only (B = m C),
// 2 stress only (zero mass, no m v, B = stress term), 3 momentum without the
// affine term (B = 0: W m v, Transfer.cu:928).
// sigf(q): entry q (row-major) of an F-based particle's stress term.
template <int STRIDE, int PART = 0, class SigF>
__device__ __forceinline__ void p2g_payload_f(const P2GIn &in, uint32_t i, const DevMaterial *mats, SigF sigf,
                                              float dx, float sdt, const float fr[3], float *pl) {
    const DevMaterial mt = *reinterpret_cast<const DevMaterial *>(reinterpret_cast<const char *>(mats) +
                                                                  (size_t)((in.mw & kMatIndexMask) *
                                                                           (uint32_t)sizeof(DevMaterial)));
    const float s = sdt * (in.pp ? chan(in.pp, i, kChVol) : mt.vol);
    float contrib[3][3];
    if (mt.model != ZMPM_MODEL_JFLUID) {
#pragma unroll
        for (int q = 0; q < 9; ++q) contrib[q / 3][q % 3] = sigf(q);
    } else {
        const float tau = mt.bulk * in.J * (in.J - 1.0f);
#pragma unroll
        for (int rr = 0; rr < 3; ++rr)
#pragma unroll
            for (int c = 0; c < 3; ++c) contrib[rr][c] = (rr == c) ? s * tau : 0.0f;
    }
    const float m = PART == 2 ? 0.0f : (in.pp ? chan(in.pp, i, kChMass) : mt.mass);
    if (PART == 1 || PART == 3) {
#pragma unroll
        for (int q = 0; q < 9; ++q) contrib[q / 3][q % 3] = 0.0f;
    }
    const float fdx[3] = {fr[0] * dx, fr[1] * dx, fr[2] * dx};
    pl[0] = m;
#pragma unroll
    for (int k = 0; k < 3; ++k) pl[(1 + k) * STRIDE] = fr[k];
#pragma unroll
    for (int rr = 0; rr < 3; ++rr) {
        float Br[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) Br[c] = PART == 3 ? 0.0f : fmaf(m, in.C[3 * rr + c], contrib[rr][c]);
        pl[(4 + rr) * STRIDE] = m * in.v[rr] - (Br[0] * fdx[0] + Br[1] * fdx[1] + Br[2] * fdx[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) pl[(7 + 3 * rr + c) * STRIDE] = Br[c] * dx;
    }
}

// The payload with the stress terms of k_stress (planes of sig_stride entries).
template <int STRIDE, int PART = 0>
__device__ __forceinline__ void p2g_payload(const P2GIn &in, uint32_t i, const DevMaterial *mats, const float *sig,
                                            int64_t sig_stride, float dx, float sdt, const float fr[3], float *pl) {
    p2g_payload_f<STRIDE, PART>(in, i, mats, [&](int q) { return sig[q * sig_stride + i]; }, dx, sdt, fr, pl);
}

// FLIP stress payload (Transfer.cu:911-929): f and S = -dt vol P F^T / dx (the
// stress term without D^-1, times the 1/dx of the weight gradients); sig holds
// k_stress's -dt D^-1 vol P F^T = S dx 4 / dx^2 (D^-1 = 4 / dx^2).
template <int STRIDE>
__device__ __forceinline__ void p2g_payload_flip_stress(const P2GIn &in, uint32_t i, const DevMaterial *mats,
                                                        const float *sig, int64_t sig_stride, float dx, float dt,
                                                        const float fr[3], float *pl) {
    const DevMaterial mt = mats[in.mw & kMatIndexMask];
    const float dxinv = 1.0f / dx;
    pl[0] = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) pl[(1 + k) * STRIDE] = fr[k];
    if (mt.model != ZMPM_MODEL_JFLUID) {
        const float back = 0.25f * dx;  // (dx^2 / 4) / dx
#pragma unroll
        for (int q = 0; q < 9; ++q) pl[(4 + q) * STRIDE] = sig[q * sig_stride + i] * back;
    } else {
        const float tau = mt.bulk * in.J * (in.J - 1.0f);
        const float d = ((-dt * (in.pp ? chan(in.pp, i, kChVol) : mt.vol)) * tau) * dxinv;
#pragma unroll
        for (int q = 0; q < 9; ++q) pl[(4 + q) * STRIDE] = (q % 4 == 0) ? d : 0.0f;
    }
}

// Phase B of the FLIP stress for x-offset ox: node (ox, oy, oz) receives
// W (S g), g = (dwx wy wz, wx dwy wz, wx wy dwz) the weight gradient per cell
// (dw0 = f - 3/2, dw1 = 2 (1 - f), dw2 = f - 1/2), into the momentum channels;
// the reference multiplies the gradient term by W once more (Transfer.cu:929).
template <int STRIDE = kP2GChunk>
__device__ __forceinline__ void p2g_accumulate_grad(float (&acc)[36], const float *s_pl, int sb, int se, int ox) {
    for (int s = sb; s < se; ++s) {
        float f[3], S[9];
#pragma unroll
        for (int q = 0; q < 3; ++q) f[q] = s_pl[(1 + q) * STRIDE + s];
#pragma unroll
        for (int q = 0; q < 9; ++q) S[q] = s_pl[(4 + q) * STRIDE + s];
        float wx[3], wy[3], wz[3];
        bspline_w(f[0], wx);
        bspline_w(f[1], wy);
        bspline_w(f[2], wz);
        const float dwy[3] = {f[1] - 1.5f, 2.0f * (1.0f - f[1]), f[1] - 0.5f};
        const float dwz[3] = {f[2] - 1.5f, 2.0f * (1.0f - f[2]), f[2] - 0.5f};
        const float wxo = ox == 0 ? wx[0] : (ox == 1 ? wx[1] : wx[2]);
        const float dwxo = ox == 0 ? f[0] - 1.5f : (ox == 1 ? 2.0f * (1.0f - f[0]) : f[0] - 0.5f);
#pragma unroll
        for (int oy = 0; oy < 3; ++oy)
#pragma unroll
            for (int oz = 0; oz < 3; ++oz) {
                const float W = wxo * wy[oy] * wz[oz];
                const float g[3] = {dwxo * wy[oy] * wz[oz], wxo * dwy[oy] * wz[oz], wxo * wy[oy] * dwz[oz]};
                const int j = 3 * oy + oz;
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    const float val = S[3 * r] * g[0] + S[3 * r + 1] * g[1] + S[3 * r + 2] * g[2];
                    acc[4 * j + 1 + r] = fmaf(W, val, acc[4 * j + 1 + r]);
                }
            }
    }
}

// Phases B to D of one chunk of a block's particles (the payload staged in LDS
// in cell order, or at the positions s_sorted lists in cell order).
// Phase B: thread = (cell, x-offset) task; each thread walks its cell's
// particles and accumulates 9 nodes x 4 channels in registers.  Phase C: the
// registers go into the x-offset's 6^3 LDS tile (the three tiles reuse the
// payload's LDS, dead by then).  Phase D: the three tiles summed in a fixed
// order, (x0 + x1) + x2, into the block's tile in HBM; a later chunk of the
// same block adds its sum to what the same thread wrote there for the earlier
// chunks (first: this is the block's first chunk), so every node's value is
// ((chunk 0 + chunk 1) + ...) whichever kernel ran the chunks.  The registers
// live only in phases B and C.  Ends with a workgroup barrier.
template <int MODE>
__device__ __forceinline__ void p2g_phase_b(float (&acc)[36], const float *s_pl, const uint8_t *s_task,
                                            const uint32_t *s_cstart, const uint16_t *s_sorted) {
    const int t = threadIdx.x;
    const int tcell = s_task[t], tox = t % 3;
    if (MODE == kP2GFlipStress)
        p2g_accumulate_grad(acc, s_pl, (int)s_cstart[tcell], (int)s_cstart[tcell + 1], tox);
    else if (s_sorted)
        p2g_accumulate(acc, s_pl, (int)s_cstart[tcell], (int)s_cstart[tcell + 1], tox, s_sorted);
    else
        p2g_accumulate(acc, s_pl, (int)s_cstart[tcell], (int)s_cstart[tcell + 1], tox);
}

// Phases C and D after the block's (or the chunk's) phase B; s_pl's payload must be dead.
__device__ __forceinline__ void p2g_tiles_out(const float (&acc)[36], float *s_pl, const uint8_t *s_task, float *tb,
                                              bool first) {
    const int t = threadIdx.x;
    // ---- phase C: registers -> this x-offset's tile (zeroed first; the payload is dead)
    float4 *s_tile = reinterpret_cast<float4 *>(s_pl);
    for (int k = t; k < 3 * kTileNodes; k += kP2GThreads) s_tile[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    __syncthreads();
    {
        const int tcell = s_task[t], tox = t % 3;
        p2g_tile_add<true>(acc, s_tile + tox * kTileNodes, tcell >> 4, (tcell >> 2) & 3, tcell & 3, tox);
    }
    // ---- phase D: the three tiles in a fixed order -> HBM (channel-major tile, tile_slot order)
    for (int k = t; k < kTileNodes; k += kP2GThreads) {
        const float4 a0 = s_tile[k], a1 = s_tile[kTileNodes + k], a2 = s_tile[2 * kTileNodes + k];
        float v[4] = {(a0.x + a1.x) + a2.x, (a0.y + a1.y) + a2.y, (a0.z + a1.z) + a2.z, (a0.w + a1.w) + a2.w};
        const int slot = tile_slot_of(k);
        if (!first) {
#pragma unroll
            for (int c = 0; c < 4; ++c) v[c] = tb[c * kTileNodes + slot] + v[c];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) tb[c * kTileNodes + slot] = v[c];
    }
    __syncthreads();
}

template <int MODE>
__device__ __forceinline__ void p2g_chunk_tail(float *s_pl, const uint8_t *s_task, const uint32_t *s_cstart,
                                               const uint16_t *s_sorted, float *tb, bool first) {
    float acc[36];
#pragma unroll
    for (int q = 0; q < 36; ++q) acc[q] = 0.0f;
    p2g_phase_b<MODE>(acc, s_pl, s_task, s_cstart, s_sorted);
    __syncthreads();
    p2g_tiles_out(acc, s_pl, s_task, tb, first);
}

// P2G (Transfer.cu:844-890) of one particle block per workgroup, into the block's
// 6^3 tile.  A block's particles are staged in chunks of kP2GChunk; the
// registers carry each lane's sums from one chunk to the next.
// Phase A1: one lane per particle finds its cell and its stable rank in the
// cell (6 wave ballots; counts per (round, wave, cell) in LDS; one barrier).
// Phase A2: the stress (fixed corotated: FP64 polar; J-fluid: tau = K J (J-1) I)
// and the 16-float affine payload, stored in cell order.  Phase B: lane = cell,
// wave = x-offset; each lane walks its cell's particles and accumulates 9
// nodes x 4 channels in registers.  Phase C: each wave adds its registers into
// a private LDS tile (the three tiles reuse the payload's LDS, which keeps the
// workgroup at 26 KB: six per CU at 96 VGPRs).  Phase D: the three tiles are summed in
// a fixed order and written to HBM.  No atomics: deterministic.
// MODE: kP2GForward (the transfer; a fixed-corotated particle's stress term
// -dt D^-1 vol P F^T comes from k_stress in sig) or kP2GAdjoint (the payload is the
// G2P adjoint's per-particle (a, B): zero mass channel, momentum channels
// W (a + B d), so the tiles receive the grid-velocity adjoint); the FLIP
// transfers' kP2GMomentum / kP2GStress (affine) and kP2GFlipMomentum /
// kP2GFlipStress (W m v, and the weight-gradient stress term).
// amdgpu_waves_per_eu(5): at most 96 VGPRs, so six workgroups fit on a CU (with 26 KB of LDS).
// PP: the particles carry their own mass and volume (L.pp_mass), a separate
// instantiation so that the common case keeps its registers.
// FLUSH: the chunk-flush grouping of the fused pass (each chunk's sums go through
// phases C and D on their own, p2g_chunk_tail); otherwise the registers carry each
// task's sums from one chunk to the next and phases C and D run once per block.
// The two agree bit for bit on blocks of at most kP2GChunk particles.
template <int MODE, bool PP = false, bool FLUSH = false>
__global__ void __launch_bounds__(kP2GThreads) __attribute__((amdgpu_waves_per_eu(5)))
k_p2g(State A, Partition P, Lattice L, const DevMaterial *mats,
                                                     const float *sig, int64_t sig_stride, float *tiles, DevStatus *st,
                                                     AdjScatter adj, uint32_t skip_upto) {
    __shared__ float s_pl[kPL * kP2GChunk];
    static_assert(3 * kTileNodes * 4 <= kPL * kP2GChunk, "tiles alias the payload");
    __shared__ uint32_t s_wcnt[kP2GRounds * 3][64];
    __shared__ uint32_t s_cstart[65];
    __shared__ uint8_t s_task[3 * 64];  // phase B task p -> cell (x-offset p % 3)

    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t nblocks = P.counts[0], nx = P.counts[2];
    const PartView V = part_view(P, L.part_shift);
    bool mismatch = false;
    // the thread's phase B task (cell s_task[t], x-offset t % 3) is read from LDS where it is
    // used rather than held in registers across the block; valid before the first sort
    s_task[t] = (uint8_t)(t / 3);

    // one part per iteration, into its own tile
    for (uint32_t j = blockIdx.x; j < nblocks + nx; j += gridDim.x) {
        const PartRef pr = part_ref(V, j, nblocks);
        const uint32_t b = pr.b;
        // k_p2g_sparse takes the blocks of at most skip_upto particles (a single part)
        if (pr.k == 0xffffffffu && pr.hi - pr.lo <= skip_upto) continue;
        int lbx, lby, lbz;
        const SceneParams sc = scene_params(L, L.nscenes == 1 ? 0u : P.pblk_scene[b]);
        const float dx = sc.dx, dxinv = sc.dxinv;
        const float sdt = -sc.dt * sc.Dinv;  // times vol per particle: -dt * Dinv * vol (Transfer.cu:869)
        lattice_coords(sc, P.pblk_lat[b], lbx, lby, lbz);
        const int bc[3] = {(lbx + L.origin) * 4, (lby + L.origin) * 4, (lbz + L.origin) * 4};
        const uint32_t ps = pr.lo, pe = pr.hi;
        float *tb = tiles + (int64_t)pr.p * (4 * kTileNodes);
        float acc[36];  // FLUSH: unused
#pragma unroll
        for (int q = 0; q < 36; ++q) acc[q] = 0.0f;

        // contiguous chunks of the block's particles (interleaving them, so that each chunk
        // samples every cell, scattered the loads of dense blocks and was slower)
        for (uint32_t c0 = ps; c0 < pe; c0 += kP2GChunk) {
            const TileRun run = tile_run(A.ch, c0);
            // ---- phase A1: cells and stable ranks of the round's particles
            int cell[kP2GRounds];
            uint32_t rank[kP2GRounds];
#pragma unroll
            for (int r = 0; r < kP2GRounds; ++r) {
                const uint32_t i = c0 + r * kP2GThreads + t;
                const bool valid = i < pe;
                cell[r] = 0;
                if (valid) {
                    int lc[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const float X = run_chan(run, run_off(run, i), k) * dxinv;
                        const int base = cell_floor(X);
                        lc[k] = base - bc[k];
                        mismatch |= (lc[k] < 0 || lc[k] > 3);
                        lc[k] = clamp03(lc[k]);
                    }
                    cell[r] = (lc[0] * 4 + lc[1]) * 4 + lc[2];
                }
                const uint64_t match = same_key_lanes<6>((uint32_t)cell[r], valid);
                rank[r] = valid ? count_below(match) : 0xffffffffu;
                s_wcnt[r * 3 + w][lane] = 0u;
                wave_lds_order();
                if (valid && rank[r] == 0) s_wcnt[r * 3 + w][cell[r]] = (uint32_t)__popcll(match);
            }
            __syncthreads();
            if (w == 0) {
                uint32_t tot = 0;
#pragma unroll
                for (int k = 0; k < kP2GRounds * 3; ++k) tot += s_wcnt[k][lane];
                uint32_t inc = tot;
#pragma unroll
                for (int off = 1; off < 64; off <<= 1) {
                    const uint32_t y = __shfl_up(inc, off, 64);
                    if (lane >= off) inc += y;
                }
                s_cstart[lane] = inc - tot;
                if (lane == 63) s_cstart[64] = inc;
                if (c0 == ps) {
                    // the 192 (cell, x-offset) tasks of phase B sorted by the cell's particle
                    // count, descending (ties by cell), so each wave gets cells of similar
                    // counts and its loop runs close to its own lanes' work
                    const uint32_t key = min(tot, 1023u);
                    uint64_t eq = ~0ull;
                    uint32_t gt = 0;
#pragma unroll
                    for (int bit = 9; bit >= 0; --bit) {
                        const bool on = (key >> bit) & 1;
                        const uint64_t bb = __ballot(on);
                        if (!on) gt += (uint32_t)__popcll(eq & bb);
                        eq &= on ? bb : ~bb;
                    }
                    const uint32_t rk = gt + count_below(eq);
#pragma unroll
                    for (int k = 0; k < 3; ++k) s_task[3 * rk + k] = (uint8_t)lane;
                }
            }
            __syncthreads();
            // ---- phase A2: stress and payload, stored in cell order
#pragma unroll
            for (int r = 0; r < kP2GRounds; ++r) {
                const uint32_t i = c0 + r * kP2GThreads + t;
                if (rank[r] == 0xffffffffu) continue;
                uint32_t slot = s_cstart[cell[r]] + rank[r];
                for (int k = 0; k < r * 3 + w; ++k) slot += s_wcnt[k][cell[r]];
                // the fractional position again, as phase A1 computed it (fewer registers held)
                float fr[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const float X = run_chan(run, run_off(run, i), k) * dxinv;
                    fr[k] = X - (float)cell_floor(X);
                }
                if (MODE == kP2GAdjoint) {
                    const float fdx[3] = {fr[0] * dx, fr[1] * dx, fr[2] * dx};
                    s_pl[slot] = 0.0f;
#pragma unroll
                    for (int k = 0; k < 3; ++k) s_pl[(1 + k) * kP2GChunk + slot] = fr[k];
#pragma unroll
                    for (int rr = 0; rr < 3; ++rr) {
                        float Br[3];
#pragma unroll
                        for (int c = 0; c < 3; ++c) Br[c] = adj.B[(3 * rr + c) * adj.stride + i];
                        s_pl[(4 + rr) * kP2GChunk + slot] =
                            adj.a[rr * adj.stride + i] - (Br[0] * fdx[0] + Br[1] * fdx[1] + Br[2] * fdx[2]);
#pragma unroll
                        for (int c = 0; c < 3; ++c) s_pl[(7 + 3 * rr + c) * kP2GChunk + slot] = Br[c] * dx;
                    }
                    continue;
                }
                P2GIn in;
                p2g_load(A, run, run_off(run, i), in, PP);
                if (MODE == kP2GFlipStress)
                    p2g_payload_flip_stress<kP2GChunk>(in, i, mats, sig, sig_stride, dx, sc.dt, fr, s_pl + slot);
                else
                    p2g_payload_f<kP2GChunk, MODE == kP2GMomentum       ? 1
                                             : MODE == kP2GStress       ? 2
                                             : MODE == kP2GFlipMomentum ? 3
                                                                        : 0>(
                        in, i, mats,
                        [&](int q) {  // sig[q * sig_stride + i] from a uniform base at the chunk
                            return *reinterpret_cast<const float *>(
                                reinterpret_cast<const char *>(sig + (size_t)q * sig_stride + c0) +
                                (size_t)((i - c0) * 4u));
                        },
                        dx, sdt, fr, s_pl + slot);
            }
            __syncthreads();
            // ---- phase B: thread = (cell, x-offset) task
            if (FLUSH) {
                p2g_chunk_tail<MODE>(s_pl, s_task, s_cstart, nullptr, tb, c0 == ps);
            } else {
                p2g_phase_b<MODE>(acc, s_pl, s_task, s_cstart, nullptr);
                __syncthreads();
            }
        }
        if (!FLUSH) p2g_tiles_out(acc, s_pl, s_task, tb, true);
    }
    if (mismatch) atomicOr(&st->bin_mismatch, 1u);
}

// P2G of the particle blocks holding at most 64 particles (kSparseMax), one wave
// per block and four blocks per workgroup, so the sparse blocks of many small
// scenes pay no workgroup barriers and leave no wave idle.  The same steps as
// k_p2g in one wave: lane = particle (cell, stable in-cell rank by 6 ballots,
// payload in cell order), then lane = cell for the three x-offsets in turn,
// each accumulated in registers and added to its own 6^3 LDS tile with ordered
// read-modify-writes (the third tile reuses the payload's LDS, dead by then),
// then the tiles summed as (x0 + x1) + x2 to HBM.  Deterministic, and the same
// floating-point operations in the same order as k_p2g for such a block, so
// the two kernels give bit-identical tiles.
constexpr int kSparseMax = 64;
__global__ void __launch_bounds__(256) k_p2g_sparse(State A, Partition P, Lattice L, const DevMaterial *mats,
                                                    const float *sig, int64_t sig_stride, float *tiles, DevStatus *st) {
    __shared__ float s_pl4[4][kPL * kSparseMax];
    static_assert(kTileNodes * 4 <= kPL * kSparseMax, "the third tile aliases the payload");
    __shared__ float4 s_tile4[4][2][kTileNodes];
    __shared__ uint32_t s_cnt4[4][65];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    float *s_pl = s_pl4[w];
    float4 *s_t0 = s_tile4[w][0], *s_t1 = s_tile4[w][1];
    float4 *s_t2 = reinterpret_cast<float4 *>(s_pl);
    uint32_t *s_cnt = s_cnt4[w];
    const int clx = lane >> 4, cly = (lane >> 2) & 3, clz = lane & 3;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t nblocks = P.counts[0];
    bool mismatch = false;
    for (uint32_t b = blockIdx.x * 4 + w; b < nblocks; b += gridDim.x * 4) {
        const uint32_t ps = P.pblk_start[b], pe = P.pblk_start[b + 1];
        if (pe - ps > (uint32_t)kSparseMax) continue;  // k_p2g takes it
        const SceneParams sc = scene_params(L, L.nscenes == 1 ? 0u : P.pblk_scene[b]);
        const float dx = sc.dx, dxinv = sc.dxinv;
        const float sdt = -sc.dt * sc.Dinv;
        int lbx, lby, lbz;
        lattice_coords(sc, P.pblk_lat[b], lbx, lby, lbz);
        const int bc[3] = {(lbx + L.origin) * 4, (lby + L.origin) * 4, (lbz + L.origin) * 4};
        // lane = particle: cell and stable rank in the cell
        const uint32_t i = ps + lane;
        const bool valid = i < pe;
        int cell = 0;
        float fr[3] = {0.0f, 0.0f, 0.0f};
        if (valid) {
            int lc[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float X = chan(A.ch, i, k) * dxinv;
                const int base = cell_floor(X);
                fr[k] = X - (float)base;
                lc[k] = base - bc[k];
                mismatch |= (lc[k] < 0 || lc[k] > 3);
                lc[k] = clamp03(lc[k]);
            }
            cell = (lc[0] * 4 + lc[1]) * 4 + lc[2];
        }
        const uint64_t match = same_key_lanes<6>((uint32_t)cell, valid);
        const uint32_t rank = (uint32_t)__popcll(match & lt_mask);
        s_cnt[lane] = 0u;
        wave_lds_order();
        if (valid && rank == 0) s_cnt[cell] = (uint32_t)__popcll(match);
        wave_lds_order();
        // first slot of every cell (lane = cell)
        const uint32_t cnt = s_cnt[lane];
        uint32_t inc = cnt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(inc, off, 64);
            if (lane >= off) inc += y;
        }
        const uint32_t sb = inc - cnt, se = inc;
        const uint32_t start = __shfl(sb, cell, 64);  // all lanes: every source lane must be active
        if (valid) {
            P2GIn in;
            p2g_load(A, i, in, L.pp_mass != 0);
            p2g_payload<kSparseMax>(in, i, mats, sig, sig_stride, dx, sdt, fr, s_pl + start + rank);
        }
        for (int k = lane; k < kTileNodes; k += 64) {
            s_t0[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            s_t1[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        wave_lds_order();
        // lane = cell, one x-offset after another, each into its own tile
        {
            float acc[36];
#pragma unroll
            for (int q = 0; q < 36; ++q) acc[q] = 0.0f;
            p2g_accumulate<kSparseMax>(acc, s_pl, (int)sb, (int)se, 0);
            p2g_tile_add<false>(acc, s_t0, clx, cly, clz, 0);
        }
        {
            float acc[36];
#pragma unroll
            for (int q = 0; q < 36; ++q) acc[q] = 0.0f;
            p2g_accumulate<kSparseMax>(acc, s_pl, (int)sb, (int)se, 1);
            p2g_tile_add<false>(acc, s_t1, clx, cly, clz, 1);
        }
        {
            float acc[36];
#pragma unroll
            for (int q = 0; q < 36; ++q) acc[q] = 0.0f;
            p2g_accumulate<kSparseMax>(acc, s_pl, (int)sb, (int)se, 2);
            wave_lds_order();  // the payload is dead: its LDS becomes the third tile
            for (int k = lane; k < kTileNodes; k += 64) s_t2[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            wave_lds_order();
            p2g_tile_add<false>(acc, s_t2, clx, cly, clz, 2);
        }
        float *tb = tiles + (int64_t)b * (4 * kTileNodes);
        for (int k = lane; k < kTileNodes; k += 64) {
            const float4 a0 = s_t0[k], a1 = s_t1[k], a2 = s_t2[k];
            const int slot = tile_slot_of(k);
            tb[slot] = (a0.x + a1.x) + a2.x;
            tb[kTileNodes + slot] = (a0.y + a1.y) + a2.y;
            tb[2 * kTileNodes + slot] = (a0.z + a1.z) + a2.z;
            tb[3 * kTileNodes + slot] = (a0.w + a1.w) + a2.w;
        }
        wave_lds_order();
    }
    if (mismatch) atomicOr(&st->bin_mismatch, 1u);
}

// ------------------------------------------------------------------ colliders
// Signed distance of a collider's shape at X (collider frame) and its gradient.
__device__ __forceinline__ float collider_phi(const DevCollider &c, const float X[3], float g[3]) {
    const float *p = c.shape_params;
    if (c.shape == ZMPM_SHAPE_HALFSPACE) {
        g[0] = p[0];
        g[1] = p[1];
        g[2] = p[2];
        return p[0] * X[0] + p[1] * X[1] + p[2] * X[2] - p[3];
    }
    if (c.shape == ZMPM_SHAPE_SPHERE) {
        const float d[3] = {X[0] - p[0], X[1] - p[1], X[2] - p[2]};
        const float r = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const float inv = r > 0.0f ? 1.0f / r : 0.0f;
        g[0] = d[0] * inv;
        g[1] = r > 0.0f ? d[1] * inv : 1.0f;
        g[2] = d[2] * inv;
        return r - p[3];
    }
    if (c.shape == ZMPM_SHAPE_BOX) {
        float q[3], sg[3], out2 = 0.0f, qmax = -3.4e38f;
        int kmax = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float cen = 0.5f * (p[k] + p[3 + k]), half = 0.5f * (p[3 + k] - p[k]);
            const float d = X[k] - cen;
            sg[k] = d < 0.0f ? -1.0f : 1.0f;
            q[k] = fabsf(d) - half;
            out2 += q[k] > 0.0f ? q[k] * q[k] : 0.0f;
            if (q[k] > qmax) {
                qmax = q[k];
                kmax = k;
            }
        }
        if (qmax > 0.0f) {  // outside: distance to the nearest point of the box
            const float o = sqrtf(out2);
#pragma unroll
            for (int k = 0; k < 3; ++k) g[k] = q[k] > 0.0f ? sg[k] * q[k] / o : 0.0f;
            return o;
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) g[k] = k == kmax ? sg[k] : 0.0f;
        return qmax;
    }
    // sampled SDF: trilinear value and the gradient of the trilinear interpolant
    float u[3], f[3];
    int i0[3];
    float s[2][2][2];
    if (c.shape == ZMPM_SHAPE_SDF_LEAVES) {
        // 8^3 leaves: each corner's leaf from the dense map over the leaves' bounding box;
        // a corner outside every leaf reads the background
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            u[k] = (X[k] - c.origin[k]) / c.spacing;
            if (!(fabsf(u[k]) < 1e9f)) {
                g[0] = g[1] = g[2] = 0.0f;
                return p[0];
            }
            const float fl = floorf(u[k]);
            i0[k] = (int)fl;
            f[k] = u[k] - fl;
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const int i = i0[0] + a, j = i0[1] + b, k = i0[2] + e;
                    const int lx = (i >> 3) - c.leaf_lo[0], ly = (j >> 3) - c.leaf_lo[1], lz = (k >> 3) - c.leaf_lo[2];
                    int32_t l = -1;
                    if ((unsigned)lx < (unsigned)c.dims[0] && (unsigned)ly < (unsigned)c.dims[1] &&
                        (unsigned)lz < (unsigned)c.dims[2])
                        l = c.leaf_map[((int64_t)lx * c.dims[1] + ly) * c.dims[2] + lz];
                    s[a][b][e] = l < 0 ? p[0] : c.sdf[(int64_t)l * 512 + ((i & 7) << 6) + ((j & 7) << 3) + (k & 7)];
                }
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            u[k] = (X[k] - c.origin[k]) / c.spacing;
            if (!(u[k] >= 0.0f && u[k] <= (float)(c.dims[k] - 1))) {
                g[0] = g[1] = g[2] = 0.0f;
                return __builtin_inff();
            }
            i0[k] = min((int)u[k], c.dims[k] - 2);
            f[k] = u[k] - (float)i0[k];
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 2; ++e)
                    s[a][b][e] = c.sdf[((int64_t)(i0[0] + a) * c.dims[1] + (i0[1] + b)) * c.dims[2] + (i0[2] + e)];
    }
    float phi = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const float wx = a ? f[0] : 1.0f - f[0], wy = b ? f[1] : 1.0f - f[1], wz = e ? f[2] : 1.0f - f[2];
                const float v = s[a][b][e];
                phi += wx * wy * wz * v;
                gx += (a ? 1.0f : -1.0f) * wy * wz * v;
                gy += wx * (b ? 1.0f : -1.0f) * wz * v;
                gz += wx * wy * (e ? 1.0f : -1.0f) * v;
            }
    g[0] = gx / c.spacing;
    g[1] = gy / c.spacing;
    g[2] = gz / c.spacing;
    return phi;
}

// zpc Collider::resolveCollision as ApplyGridBoundaryOnZSGrid spells out the
// response (Boundary.cu:55-66), on node position x and velocity v.
__device__ __forceinline__ void resolve_collision(const DevCollider &c, const float x[3], float v[3]) {
    const float d[3] = {x[0] - c.b[0], x[1] - c.b[1], x[2] - c.b[2]};
    const float *R = c.R;
    const float inv_s = 1.0f / c.s;
    float X[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) X[k] = (R[k] * d[0] + R[3 + k] * d[1] + R[6 + k] * d[2]) * inv_s;
    float g[3];
    const float phi = collider_phi(c, X, g) * c.s;
    if (!(phi < 0.0f)) return;
    const float rs = c.dsdt * inv_s;
    float vo[3];
    vo[0] = c.dbdt[0] + (c.omega[1] * d[2] - c.omega[2] * d[1]) + rs * d[0];
    vo[1] = c.dbdt[1] + (c.omega[2] * d[0] - c.omega[0] * d[2]) + rs * d[1];
    vo[2] = c.dbdt[2] + (c.omega[0] * d[1] - c.omega[1] * d[0]) + rs * d[2];
    if (c.type == ZMPM_COLLIDER_STICKY) {
        v[0] = vo[0];
        v[1] = vo[1];
        v[2] = vo[2];
        return;
    }
    float n[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) n[k] = R[3 * k] * g[0] + R[3 * k + 1] * g[1] + R[3 * k + 2] * g[2];
    const float nn = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (!(nn > 0.0f)) return;
    float vr[3] = {v[0] - vo[0], v[1] - vo[1], v[2] - vo[2]};
    const float proj = (n[0] * vr[0] + n[1] * vr[1] + n[2] * vr[2]) / (nn * nn);
    if (c.type == ZMPM_COLLIDER_SLIP || proj < 0.0f) {
#pragma unroll
        for (int k = 0; k < 3; ++k) vr[k] -= proj * n[k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] = vr[k] + vo[k];
}

// Level-set colliders on the updated grid, one wave per grid block, lane = node
// with m != 0, at the node position (block key + cell) dx (Boundary.cu:108-113).
// vmax: a moving collider can speed a node up, so the block's maximum speed
// (which bounds the move pass) takes in the final velocities.
__global__ void __launch_bounds__(kThreads) k_colliders(Partition P, Lattice L, float *grid, float *gvmax, int vmax) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lx = lane >> 4, ly = (lane >> 2) & 3, lz = lane & 3;
    const uint32_t ng = P.counts[1];
    for (uint32_t gb = blockIdx.x * 4 + w; gb < ng; gb += gridDim.x * 4) {
        const uint32_t scene = L.nscenes == 1 ? 0u : P.gblk_scene[gb];
        const SceneParams sc = scene_params(L, scene);
        int Lx, Ly, Lz;
        lattice_coords(sc, P.gblk_lat[gb], Lx, Ly, Lz);
        float *blk = grid + (int64_t)gb * 256;
        float v[3] = {blk[64 + lane], blk[128 + lane], blk[192 + lane]};
        if (blk[lane] != 0.0f) {
            const float pos[3] = {(float)((Lx + L.origin) * 4 + lx) * sc.dx, (float)((Ly + L.origin) * 4 + ly) * sc.dx,
                                  (float)((Lz + L.origin) * 4 + lz) * sc.dx};
            for (int c = 0; c < L.ncols; ++c)
                if ((uint32_t)L.cols[c].scene == scene) resolve_collision(L.cols[c], pos, v);
            blk[64 + lane] = v[0];
            blk[128 + lane] = v[1];
            blk[192 + lane] = v[2];
        }
        if (vmax) {
            float m2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            if (!(m2 <= 3.0e38f)) m2 = __builtin_inff();
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) m2 = fmaxf(m2, __shfl_xor(m2, off, 64));
            if (lane == 0) gvmax[gb] = fmaxf(gvmax[gb], m2);
        }
    }
}

// ------------------------------------------------------------------ boundary particles
// ZSBoundaryPrimitiveToZSGrid's p2g_momentum (Transfer.cu:1533-1578) onto the
// simulation grid's blocks, then ApplyGridBoundaryOnZSGrid (Boundary.cu:18-74).
// The sums are 64-bit fixed point: integer additions commute, so the result does
// not depend on the order of the atomics.  The scale is a power of two chosen per
// substep from the largest |m| max(1, |v|, |nrm|) so that n such terms stay below
// 2^61.

// the clear of the grid blocks' sums and of the scale word
__global__ void __launch_bounds__(kThreads) k_bnd_clear(BoundaryParticles bp, Partition P) {
    const int64_t total = (int64_t)P.counts[1] * (kBndChannels * 64);
    const int64_t gtid = blockIdx.x * (int64_t)kThreads + threadIdx.x, gsz = (int64_t)gridDim.x * kThreads;
    for (int64_t k = gtid; k < total; k += gsz) bp.acc[k] = 0ull;
    if (gtid == 0) *bp.scale_bits = 0u;
}

// largest |m| max(1, |v|, |nrm|): an integer max on the bits of non-negative floats
__global__ void __launch_bounds__(kThreads) k_bnd_scale(BoundaryParticles bp) {
    float a = 0.0f;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < bp.n; i += (int64_t)gridDim.x * kThreads) {
        const float v0 = bp.v[i], v1 = bp.v[bp.n + i], v2 = bp.v[2 * bp.n + i];
        float s = fmaxf(1.0f, sqrtf(v0 * v0 + v1 * v1 + v2 * v2));
        if (bp.nrm) {
            const float n0 = bp.nrm[i], n1 = bp.nrm[bp.n + i], n2 = bp.nrm[2 * bp.n + i];
            s = fmaxf(s, sqrtf(n0 * n0 + n1 * n1 + n2 * n2));
        }
        const float t = fabsf(bp.m[i]) * s;
        a = fmaxf(a, t <= 3.0e38f ? t : 3.0e38f);  // a non-finite input saturates (and is reported)
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a = fmaxf(a, __shfl_xor(a, off, 64));
    if ((threadIdx.x & 63) == 0 && a > 0.0f) atomicMax(bp.scale_bits, __float_as_uint(a));
}

__device__ __forceinline__ double bnd_scale(const BoundaryParticles &bp) {
    const double amax = fmax((double)__uint_as_float(*bp.scale_bits), 1e-30);
    return ldexp(1.0, 61 - (int)ceil(log2((double)bp.n * amax)));
}

// one lane per boundary particle: the linear kernel (base floor(x / dx), weights
// 1 - f and f per axis) onto the 8 nodes that lie in the partition's grid blocks
__global__ void __launch_bounds__(kThreads) k_bnd_p2g(BoundaryParticles bp, Partition P, Lattice L) {
    const SceneParams sc = scene_params(L, (uint32_t)bp.scene);
    const double scale = bnd_scale(bp);
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < bp.n; i += (int64_t)gridDim.x * kThreads) {
        int base[3];
        float w[3][2];
        bool inside = true;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float X = bp.x[k * bp.n + i] * sc.dxinv;
            const double c = floor((double)X);
            inside = inside && c > -1.0e9 && c < 1.0e9;
            base[k] = inside ? (int)c : 0;
            const float f = X - (float)base[k];
            w[k][0] = 1.0f - f;
            w[k][1] = f;
        }
        if (!inside) continue;
        const double m = (double)bp.m[i];
        double val[kBndChannels];
        val[0] = m;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            val[1 + k] = m * (double)bp.v[k * bp.n + i];
            val[4 + k] = bp.nrm ? m * (double)bp.nrm[k * bp.n + i] : 0.0;
        }
#pragma unroll
        for (int o = 0; o < 8; ++o) {
            const int g[3] = {base[0] + (o >> 2), base[1] + ((o >> 1) & 1), base[2] + (o & 1)};
            const int lb[3] = {(g[0] >> 2) - L.origin, (g[1] >> 2) - L.origin, (g[2] >> 2) - L.origin};
            if (lb[0] < 0 || lb[1] < 0 || lb[2] < 0 || lb[0] >= sc.nb || lb[1] >= sc.nb || lb[2] >= sc.nb) continue;
            const int64_t l = sc.lat_off + ((int64_t)lb[0] * sc.nb + lb[1]) * sc.nb + lb[2];
            const int32_t gid = P.lat_gid[l];
            if (gid < 0) continue;  // no simulation node there: nothing to constrain
            const int cell = ((g[0] & 3) * 4 + (g[1] & 3)) * 4 + (g[2] & 3);
            const double W = (double)w[0][o >> 2] * (double)w[1][(o >> 1) & 1] * (double)w[2][o & 1];
            unsigned long long *a = bp.acc + (int64_t)gid * (kBndChannels * 64) + cell;
#pragma unroll
            for (int c = 0; c < kBndChannels; ++c)
                atomicAdd(a + c * 64, (unsigned long long)llrint(W * val[c] * scale));
        }
    }
}

// ApplyGridBoundaryOnZSGrid on every node with mass in both grids: v_b = (m v)/m,
// n_b = normalize((m nrm)/m) (the "boundary" grid update, MPMPipeline.cu:896-910);
// sticky v = v_b, else v - v_b loses its n_b part (separate: only if approaching).
// One wave per grid block, lane = node; vmax folds the final speeds into the block maxima.
__global__ void __launch_bounds__(kThreads) k_bnd_apply(BoundaryParticles bp, Partition P, Lattice L, float *grid,
                                                        float *gvmax) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t ng = P.counts[1];
    const double inv_scale = 1.0 / bnd_scale(bp);
    for (uint32_t gb = blockIdx.x * 4 + w; gb < ng; gb += gridDim.x * 4) {
        const uint32_t scene = L.nscenes == 1 ? 0u : P.gblk_scene[gb];
        if (scene != (uint32_t)bp.scene) continue;
        float *blk = grid + (int64_t)gb * 256;
        const unsigned long long *a = bp.acc + (int64_t)gb * (kBndChannels * 64) + lane;
        const double bm = (double)(long long)a[0] * inv_scale;
        float v[3] = {blk[64 + lane], blk[128 + lane], blk[192 + lane]};
        if (blk[lane] != 0.0f && bm != 0.0) {
            const double im = 1.0 / bm;
            double vb[3], nb[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                vb[k] = (double)(long long)a[(1 + k) * 64] * inv_scale * im;
                nb[k] = (double)(long long)a[(4 + k) * 64] * inv_scale * im;
            }
            if (bp.type == ZMPM_COLLIDER_STICKY) {
#pragma unroll
                for (int k = 0; k < 3; ++k) v[k] = (float)vb[k];
            } else {
                const double nn = sqrt(nb[0] * nb[0] + nb[1] * nb[1] + nb[2] * nb[2]);
                if (nn > 0.0) {
                    double r[3], proj = 0.0;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        nb[k] /= nn;
                        r[k] = (double)v[k] - vb[k];
                        proj += nb[k] * r[k];
                    }
                    if (bp.type == ZMPM_COLLIDER_SLIP || proj < 0.0) {
#pragma unroll
                        for (int k = 0; k < 3; ++k) r[k] -= proj * nb[k];
                    }
#pragma unroll
                    for (int k = 0; k < 3; ++k) v[k] = (float)(r[k] + vb[k]);
                }
            }
            blk[64 + lane] = v[0];
            blk[128 + lane] = v[1];
            blk[192 + lane] = v[2];
        }
        float m2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        if (!(m2 <= 3.0e38f)) m2 = __builtin_inff();
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m2 = fmaxf(m2, __shfl_xor(m2, off, 64));
        if (lane == 0 && gvmax) gvmax[gb] = fmaxf(gvmax[gb], m2);
    }
}

// StepZSBoundary: b += dbdt dt, s += dsdt dt, dt the collider scene's (dt < 0) or the given one.
__global__ void k_step_colliders(Lattice L, float dt) {
    const int c = threadIdx.x;
    if (c >= L.ncols) return;
    DevCollider &col = L.cols[c];
    const float h = dt >= 0.0f ? dt : scene_params(L, (uint32_t)col.scene).dt;
#pragma unroll
    for (int k = 0; k < 3; ++k) col.b[k] += col.dbdt[k] * h;
    col.s += col.dsdt * h;
}

// ------------------------------------------------------------------ grid
// The node (lane) of grid block l summed over the <= 8 particle blocks whose 6^3
// tiles overlap it, in a fixed order; lookups are unconditional (clamped) so
// they are all in flight.
__device__ __forceinline__ void combine_tiles(const Partition &P, const float *tiles, int64_t l, int64_t nb,
                                              int64_t nb2, int Lx, int Ly, int Lz, int lane, float &m, float &p0,
                                              float &p1, float &p2) {
    const int lx = lane >> 4, ly = (lane >> 2) & 3, lz = lane & 3;
    int32_t src[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int ex = e >> 2, ey = (e >> 1) & 1, ez = e & 1;
        const bool in = Lx - ex >= 0 && Ly - ey >= 0 && Lz - ez >= 0;
        const int32_t v = P.lat_pid[in ? l - ex * nb2 - ey * nb - ez : l];
        const bool covers = !((ex && lx >= 2) || (ey && ly >= 2) || (ez && lz >= 2));
        src[e] = (in && covers) ? v : -1;
    }
    float tv[8][4];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int ey = (e >> 1) & 1, ez = e & 1;
        // region e of the tile of particle block B - e: this grid block's nodes, contiguous
        const int tt = tile_region_offset(e) + (lx * (4 - 2 * ey) + ly) * (4 - 2 * ez) + lz;
        const float *tp = tiles + (int64_t)(src[e] >= 0 ? src[e] : 0) * (4 * kTileNodes) + (src[e] >= 0 ? tt : 0);
#pragma unroll
        for (int c = 0; c < 4; ++c) tv[e][c] = tp[c * kTileNodes];
    }
    m = p0 = p1 = p2 = 0.0f;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        if (src[e] < 0) continue;
        m += tv[e][0];
        p0 += tv[e][1];
        p1 += tv[e][2];
        p2 += tv[e][3];
    }
    if (P.counts[2] != 0u) {  // uniform: the extra parts of blocks cut into several, in (e, part) order
        for (int e = 0; e < 8; ++e) {
            // the source looked up again (rather than kept live through the common path)
            const int ex = e >> 2, ey = (e >> 1) & 1, ez = e & 1;
            const bool in = Lx - ex >= 0 && Ly - ey >= 0 && Lz - ez >= 0;
            const bool covers = !((ex && lx >= 2) || (ey && ly >= 2) || (ez && lz >= 2));
            if (!in || !covers || !P.lat_xf[l - ex * nb2 - ey * nb - ez]) continue;
            const int32_t sb = P.lat_pid[l - ex * nb2 - ey * nb - ez];
            const int tt = tile_region_offset(e) + (lx * (4 - 2 * ey) + ly) * (4 - 2 * ez) + lz;
            const uint32_t x1 = P.pblk_x[sb + 1];
            for (uint32_t k = P.pblk_x[sb]; k < x1; ++k) {
                const float *tp = tiles + (int64_t)(P.parts_cap - 1u - k) * (4 * kTileNodes) + tt;
                m += tp[0];
                p0 += tp[kTileNodes];
                p1 += tp[2 * kTileNodes];
                p2 += tp[3 * kTileNodes];
            }
        }
    }
}

// Grid stages of the affine FLIP transfers (UpdateZSGrid flip branch,
// MPMPipeline.cu:878-895): grid holds (m, v), gstar (m, vstar).  COMBINE sums the
// momentum tiles into grid and the stress tiles into gstar; UPDATE forms
// v = mv / m and vstar = fs / m + v + a dt; BOUNDARY (the sticky box) acts on
// vstar.  The per-block maximum takes both |v|^2 and |vstar|^2.
__global__ void __launch_bounds__(kThreads) k_grid_aflip(int mode, Partition P, Lattice L, const float *tiles,
                                                         const float *tiles2, float *grid, float *gstar,
                                                         float *gvmax, uint32_t *vclear, int32_t nclear) {
    if (vclear && blockIdx.x == 0)  // as k_grid
        for (int32_t q = threadIdx.x; q < nclear; q += kThreads) vclear[q] = 0u;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lx = lane >> 4, ly = (lane >> 2) & 3, lz = lane & 3;
    const uint32_t ng = P.counts[1];
    for (uint32_t gb = blockIdx.x * 4 + w; gb < ng; gb += gridDim.x * 4) {
        const int64_t l = P.gblk_lat[gb];
        const SceneParams sc = scene_params(L, L.nscenes == 1 ? 0u : P.gblk_scene[gb]);
        const int64_t nb = sc.nb, nb2 = nb * nb;
        int Lx, Ly, Lz;
        lattice_coords(sc, l, Lx, Ly, Lz);
        float vmax = 0.0f;
        float *blk = grid + (int64_t)gb * 256;
        float *sblk = gstar + (int64_t)gb * 256;
        float m, p[3], s[3];
        if (mode & GRID_CLEAR) {
            m = p[0] = p[1] = p[2] = s[0] = s[1] = s[2] = 0.0f;
        } else if (mode & GRID_COMBINE) {
            float m2;
            combine_tiles(P, tiles, l, nb, nb2, Lx, Ly, Lz, lane, m, p[0], p[1], p[2]);
            combine_tiles(P, tiles2, l, nb, nb2, Lx, Ly, Lz, lane, m2, s[0], s[1], s[2]);
        } else {
            m = blk[lane];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                p[k] = blk[64 * (k + 1) + lane];
                s[k] = sblk[64 * (k + 1) + lane];
            }
        }
        if ((mode & GRID_UPDATE) && m != 0.0f) {
            const float inv = 1.0f / m;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                p[k] = p[k] * inv;
                s[k] = s[k] * inv + p[k] + sc.accel[k] * sc.dt;
            }
            vmax = fmaxf(p[0] * p[0] + p[1] * p[1] + p[2] * p[2], s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        }
        if ((mode & GRID_BOUNDARY) && sc.boundary == ZMPM_BOUNDARY_STICKY_BOX && m != 0.0f) {
            const int gx = (Lx + L.origin) * 4 + lx, gy = (Ly + L.origin) * 4 + ly, gz = (Lz + L.origin) * 4 + lz;
            const int lo = sc.pad, hi = sc.n_cells - sc.pad;
            if (gx < lo || gy < lo || gz < lo || gx > hi || gy > hi || gz > hi) s[0] = s[1] = s[2] = 0.0f;
        }
        blk[lane] = m;
        sblk[lane] = m;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            blk[64 * (k + 1) + lane] = p[k];
            sblk[64 * (k + 1) + lane] = s[k];
        }
        if (mode & GRID_UPDATE) {
            if (!(vmax <= 3.0e38f)) vmax = __builtin_inff();
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off, 64));
            if (lane == 0) gvmax[gb] = vmax;
        }
    }
}

// out: where the result goes (grid itself unless GRID_KEEP, which also leaves the
// combined mass and momentum in grid and writes the updated velocities to out).
// vclear (optional): the substep's per-scene MaxVelSqr words, cleared here by the
// first workgroup for k_reduce_vmax, the next launch but the colliders.
__global__ void __launch_bounds__(kThreads) k_grid(int mode, Partition P, Lattice L, const float *tiles, float *grid,
                                                   float *gvmax, float *out, uint32_t *vclear, int32_t nclear) {
    if (vclear && blockIdx.x == 0)
        for (int32_t q = threadIdx.x; q < nclear; q += kThreads) vclear[q] = 0u;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lx = lane >> 4, ly = (lane >> 2) & 3, lz = lane & 3;
    const uint32_t ng = P.counts[1];
    for (uint32_t gb = blockIdx.x * 4 + w; gb < ng; gb += gridDim.x * 4) {
        const int64_t l = P.gblk_lat[gb];
        const SceneParams sc = scene_params(L, L.nscenes == 1 ? 0u : P.gblk_scene[gb]);
        const int64_t nb = sc.nb, nb2 = nb * nb;
        int Lx, Ly, Lz;
        lattice_coords(sc, l, Lx, Ly, Lz);
        float vmax = 0.0f;
        float *blk = grid + (int64_t)gb * 256;
        float *oblk = out + (int64_t)gb * 256;
        float m, p0, p1, p2;
        if (mode & GRID_CLEAR) {
            m = p0 = p1 = p2 = 0.0f;
        } else if (mode & GRID_COMBINE) {
            combine_tiles(P, tiles, l, nb, nb2, Lx, Ly, Lz, lane, m, p0, p1, p2);
        } else {
            m = blk[lane];
            p0 = blk[64 + lane];
            p1 = blk[128 + lane];
            p2 = blk[192 + lane];
        }
        if (mode & GRID_KEEP) {
            blk[lane] = m;
            blk[64 + lane] = p0;
            blk[128 + lane] = p1;
            blk[192 + lane] = p2;
        }
        if ((mode & GRID_UPDATE) && m != 0.0f) {
            // UpdateZSGrid apic branch (MPMPipeline.cu:857-875)
            const float inv = 1.0f / m;
            p0 = p0 * inv + sc.accel[0] * sc.dt;
            p1 = p1 * inv + sc.accel[1] * sc.dt;
            p2 = p2 * inv + sc.accel[2] * sc.dt;
            vmax = p0 * p0 + p1 * p1 + p2 * p2;
        }
        if ((mode & GRID_BOUNDARY) && sc.boundary == ZMPM_BOUNDARY_STICKY_BOX && m != 0.0f) {
            const int gx = (Lx + L.origin) * 4 + lx, gy = (Ly + L.origin) * 4 + ly, gz = (Lz + L.origin) * 4 + lz;
            const int lo = sc.pad, hi = sc.n_cells - sc.pad;
            if (gx < lo || gy < lo || gz < lo || gx > hi || gy > hi || gz > hi) p0 = p1 = p2 = 0.0f;
        }
        if ((mode & (GRID_CLEAR | GRID_COMBINE)) || out != grid) oblk[lane] = m;
        oblk[64 + lane] = p0;
        oblk[128 + lane] = p1;
        oblk[192 + lane] = p2;
        if (mode & GRID_UPDATE) {
            // max |v|^2 of the block (wave reduction), kept per grid block; the per-scene
            // maxima are reduced from these only when asked for (k_reduce_vmax)
            if (!(vmax <= 3.0e38f)) vmax = __builtin_inff();  // a non-finite velocity is not dropped by fmaxf
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off, 64));
            if (lane == 0) gvmax[gb] = vmax;
        }
    }
}

// ------------------------------------------------------------------ G2P
// Direction (0..26, 13 = stay) from block lb to the block of the advected
// position x + v dt, v interpolated from the staged grid velocities; the same
// arithmetic as the update pass, so both agree on every particle's block.
__device__ __forceinline__ uint32_t direction_of(const float xn3[3], const SceneParams &sc, const int lb[3],
                                                 int origin, bool &bad, bool &cfl) {
    int d = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float xn = xn3[k];
        const double c = (double)(xn * sc.dxinv) - 0.5;
        int nbk;
        if (c > -1.0e9 && c < 1.0e9) {
            nbk = ((int)floor(c) >> 2) - origin;
        } else {
            nbk = lb[k];
            bad = true;
        }
        if (nbk < 0 || nbk > sc.nb - 2) {
            bad = true;
            nbk = nbk < 0 ? 0 : sc.nb - 2;
        }
        int dk = nbk - lb[k];
        if (dk < -1 || dk > 1) {
            cfl = true;
            dk = dk < -1 ? -1 : 1;
        }
        d = d * 3 + (dk + 1);
    }
    return (uint32_t)d;
}

__device__ __forceinline__ uint32_t move_direction(const float4 *s_v, const float xp[3], const SceneParams &sc,
                                                   const int bc[3], const int lb[3], int origin, bool &bad, bool &cfl) {
    Arena ar;
    make_arena(xp, sc.dxinv, ar);
    const int lc[3] = {clamp03(ar.base[0] - bc[0]), clamp03(ar.base[1] - bc[1]), clamp03(ar.base[2] - bc[2])};
    float vel[3], Cn[3][3];
    g2p_gather<false>(s_v, ar, lc, sc.dx, sc.Dinv, vel, Cn);
    float xn[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) xn[k] = xp[k] + vel[k] * sc.dt;
    return direction_of(xn, sc, lb, origin, bad, cfl);
}

// AFLIP/ASFLIP particle update (Transfer.cu:1199-1206, flip = 0.99):
// v = vstar + flip (v0 - v), x += dt (vstar + beta flip (v0 - v)).
constexpr float kFlipBlend = 0.99f;
__device__ __forceinline__ void aflip_blend(const float xp[3], const float v0[3], const float vel[3],
                                            const float vst[3], float beta, float dt, float xn[3], float vn[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float dv = v0[k] - vel[k];
        vn[k] = vst[k] + kFlipBlend * dv;
        xn[k] = xp[k] + dt * (vst[k] + beta * kFlipBlend * dv);
    }
}

__device__ __forceinline__ uint32_t move_direction_aflip(const float4 *s_v, const float4 *s_vs, const float xp[3],
                                                         const float v0[3], float beta, const SceneParams &sc,
                                                         const int bc[3], const int lb[3], int origin, bool &bad,
                                                         bool &cfl) {
    Arena ar;
    make_arena(xp, sc.dxinv, ar);
    const int lc[3] = {clamp03(ar.base[0] - bc[0]), clamp03(ar.base[1] - bc[1]), clamp03(ar.base[2] - bc[2])};
    float vel[3], vst[3], Cn[3][3];
    g2p_gather<false>(s_v, ar, lc, sc.dx, sc.Dinv, vel, Cn);
    g2p_gather<false>(s_vs, ar, lc, sc.dx, sc.Dinv, vst, Cn);
    float xn[3], vn[3];
    aflip_blend(xp, v0, vel, vst, beta, sc.dt, xn, vn);
    return direction_of(xn, sc, lb, origin, bad, cfl);
}

constexpr int kMoveList = 1024;  // particle blocks up to this size use the kept/listed split
constexpr uint32_t kDirStay = 13;

// G2P pass 1 of the fused substep: for every particle, the direction (one of
// 27) from its block to the block of its advected position and its stable rank
// among the particles of its block moving that way (5 wave ballots per chunk);
// per block, the count per direction; per destination block, the incoming
// particle count and the particle/grid block marks.  Also resets the state of
// the lattice scan that follows.
// A particle's interpolated velocity is a convex combination of its 27 nodes'
// velocities, so |v| is at most the largest node speed of the 8 grid blocks
// around its particle block (k_grid's per-block maxima).  A particle farther
// than that bound times dt (plus a rounding margin) from its block's faces
// keeps its block: it gets direction 13 without the gather.  The others (a few
// per cent at CFL-limited steps) are listed and gathered together, so the
// 27-node interpolation runs on full waves; a block stages its grid velocities
// in LDS only when it lists a particle, and a block segment that lists none
// takes its ranks from the particle order without ballots.  Every direction
// equals the one the full interpolation gives.
// NT: threads per workgroup (64 or 256; picked per context by zmpm_autotune).
// AFLIP: the FLIP transfers (grid velocities v and vstar, the particle's own v
// blended in, the particle's beta or 0 for non-affine FLIP); every particle is
// interpolated.
// amdgpu_waves_per_eu(5): at most 96 VGPRs (the compiler then needs about 70), so more
// particle blocks are in flight per CU; the interpolation is rare.
// FUSED (for the fused G2P / P2G pass, APIC): in place of the codes, the block's
// stayers (direction 13) are listed in order at B.code[ps + rank] and its movers
// at B.mvlist[ps + k] (k-th mover of the block in particle order) as
// (index << 5) | direction, which the destination blocks filter by direction.
template <int NT, bool AFLIP = false, bool FUSED = false>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(5)))
k_g2p_move(State A, Partition P, Lattice L, const float *grid,
                                                 const float *gvmax, Binning B, DevStatus *st, const float *gstar,
                                                 int nonaffine) {
    constexpr int NW = NT / 64;
    __shared__ float4 s_v[kTileNodes];
    __shared__ float4 s_vs[AFLIP ? kTileNodes : 1];
    __shared__ uint32_t s_run[32];
    __shared__ uint32_t s_wcnt[NW][32];
    __shared__ uint32_t s_mcnt[NW];
    __shared__ uint32_t s_mrun;
    __shared__ uint8_t s_pdir[kMoveList];    // direction of each particle of the block (kept/listed split)
    __shared__ uint16_t s_list[kMoveList];   // offsets of the particles that need the interpolation
    __shared__ uint32_t s_nlist;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    {
        const int64_t gtid = blockIdx.x * (int64_t)NT + t, gsz = (int64_t)gridDim.x * NT;
        for (int64_t k = gtid; k < B.scan_tiles; k += gsz) {
            B.scan_p.status[k] = 0ull;
            B.scan_g.status[k] = 0ull;
        }
        if (gtid == 0) *B.scan_p.tile_ctr = 0u;
    }
    // UpdateZSGrid's MaxVelSqr of this substep (MPMPipeline.cu:939-942): each scene's maximum of
    // the grid blocks' maxima, after the colliders and boundary particles, into the words k_grid
    // cleared (one integer atomic max on the float bits per wave of one scene, else per lane;
    // order free, so deterministic)
    if (B.vmax_out) {
        const uint32_t ng = P.counts[1];
        for (int64_t g0 = blockIdx.x * (int64_t)NT; g0 < ng; g0 += (int64_t)gridDim.x * NT) {  // uniform
            const int64_t g = g0 + t;
            const bool valid = g < ng;
            const uint32_t scene = valid ? (L.nscenes == 1 ? 0u : P.gblk_scene[g]) : 0xffffffffu;
            float v = valid ? gvmax[g] : 0.0f;
            const uint32_t s0 = __shfl(scene, 0, 64);
            if (__ballot(valid && scene != s0) == 0ull) {
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
                if (lane == 0 && v > 0.0f) atomicMax(&B.vmax_out[s0], __float_as_uint(v));
            } else if (valid && v > 0.0f) {
                atomicMax(&B.vmax_out[scene], __float_as_uint(v));
            }
        }
    }
    const uint32_t nblocks = P.counts[0], nx = P.counts[2];
    const uint32_t nunits = nblocks + nx;
    const PartView V = part_view(P, L.part_shift);
    bool bad = false, cfl = false;
    // one part per iteration (FUSED: the partition has one part per block), in XCD slab order;
    // ranks count within the part
    for (uint32_t j = xcd_first(nunits); j < nunits; j += gridDim.x) {
        const PartRef pr = part_ref(V, j, nblocks);
        const uint32_t b = pr.b;
        const int64_t l = P.pblk_lat[b];
        const uint32_t ps = pr.lo, pe = pr.hi;
        const SceneParams sc = scene_params(L, L.nscenes == 1 ? 0u : P.pblk_scene[b]);
        const int64_t nb = sc.nb, nb2 = nb * nb;
        float xq[2][3];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const uint32_t i = min(ps + q * NT + t, pe - 1);
#pragma unroll
            for (int k = 0; k < 3; ++k) xq[q][k] = chan(A.ch, i, kChX + k);
        }
        int lb[3];
        lattice_coords(sc, l, lb[0], lb[1], lb[2]);
        const int bc[3] = {(lb[0] + L.origin) * 4, (lb[1] + L.origin) * 4, (lb[2] + L.origin) * 4};
        // speed bound over the stencils of this block's particles (grid blocks lb + {0,1}^3)
        float v2 = 0.0f;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int32_t gid = P.lat_gid[l + (e >> 2) * nb2 + ((e >> 1) & 1) * nb + (e & 1)];
            v2 = fmaxf(v2, gid >= 0 ? gvmax[gid] : __builtin_inff());
        }
        const float reach = AFLIP ? __builtin_inff() : sqrtf(v2) * (sc.dt * sc.dxinv) * 1.001f;  // cells (upper bound)
        bool staged = false;  // grid velocities in LDS: only once a particle needs them
        if (t < 32) s_run[t] = 0u;
        if (t == 0) {
            s_nlist = 0u;
            s_mrun = 0u;
        }
        __syncthreads();
        // segments of up to kMoveList particles: kept particles get direction 13, the
        // others are listed (order is irrelevant: each listed particle only writes its
        // own direction) and interpolated together
        for (uint32_t s0 = ps; s0 < pe; s0 += kMoveList) {
            const uint32_t s1 = min(pe, s0 + (uint32_t)kMoveList);
            for (uint32_t c0 = s0; c0 < s1; c0 += NT) {
                const uint32_t q = (c0 - ps) / NT;
                const uint32_t i = c0 + t;
                const bool valid = i < s1;
                float xp[3];
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    xp[k] = q == 0 ? xq[0][k] : (q == 1 ? xq[1][k] : chan(A.ch, min(i, pe - 1), kChX + k));
                bool keep = valid;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const float X = xp[k] * sc.dxinv;
                    const float d = (X - 0.5f) - (float)bc[k];  // cells to the lower face
                    const float margin = reach + (fabsf(X) + 4.0f) * 2.0e-6f;
                    keep = keep && d > margin && 4.0f - d > margin;  // false for NaN
                }
                if (keep) s_pdir[i - s0] = (uint8_t)kDirStay;
                const uint64_t lm = __ballot(valid && !keep);
                if (lm != 0ull) {
                    uint32_t base = 0u;
                    if (lane == 0) base = atomicAdd(&s_nlist, (uint32_t)__popcll(lm));
                    base = __shfl(base, 0, 64);
                    if (valid && !keep) s_list[base + (uint32_t)__popcll(lm & lt_mask)] = (uint16_t)(i - s0);
                }
            }
            __syncthreads();
            const uint32_t nl = s_nlist;
            if (nl > 0u && !staged) {  // uniform: nl was read after the barrier
                stage_block_velocity<NT>(s_v, grid, P.lat_gid, l, nb, nb2);
                if (AFLIP) stage_block_velocity<NT>(s_vs, gstar, P.lat_gid, l, nb, nb2);
                staged = true;
                __syncthreads();
            }
#pragma unroll 1
            for (uint32_t j = t; j < nl; j += NT) {
                const uint32_t off = s_list[j];
                float xp[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) xp[k] = chan(A.ch, s0 + off, kChX + k);
                if (AFLIP) {
                    float v0[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k) v0[k] = chan(A.ch, s0 + off, kChV + k);
                    const float beta = nonaffine ? 0.0f : chan(A.ch, s0 + off, kChBeta);
                    s_pdir[off] =
                        (uint8_t)move_direction_aflip(s_v, s_vs, xp, v0, beta, sc, bc, lb, L.origin, bad, cfl);
                } else {
                    s_pdir[off] = (uint8_t)move_direction(s_v, xp, sc, bc, lb, L.origin, bad, cfl);
                }
            }
            __syncthreads();
            if (t == 0) s_nlist = 0u;  // the next segment's list starts after the barriers below
            if (nl == 0u) {
                // every particle of the segment keeps its block: its rank is its place in the block
                const uint32_t r0 = s_run[kDirStay];
                for (uint32_t i = s0 + t; i < s1; i += NT) {
                    if (FUSED) B.code[ps + r0 + (i - s0)] = i;
                    else B.code[i] = ((r0 + (i - s0)) << 5) | kDirStay;
                }
                __syncthreads();
                if (t == 0) s_run[kDirStay] = r0 + (s1 - s0);
                __syncthreads();
                continue;
            }
            for (uint32_t c0 = s0; c0 < s1; c0 += NT) {
                const uint32_t i = c0 + t;
                const bool valid = i < s1;
                const uint32_t dir = valid ? (uint32_t)s_pdir[i - s0] : 31u;
                if (t < NW * 32) s_wcnt[t >> 5][t & 31] = 0u;
                __syncthreads();
                const uint64_t match = same_key_lanes<5>(dir, valid);
                const uint32_t rank_w = (uint32_t)__popcll(match & lt_mask);
                if (valid && rank_w == 0) s_wcnt[w][dir] = (uint32_t)__popcll(match);
                const uint64_t mm = FUSED ? __ballot(valid && dir != kDirStay) : 0ull;
                if (FUSED && lane == 0) s_mcnt[w] = (uint32_t)__popcll(mm);
                __syncthreads();
                if (valid) {
                    uint32_t rank = s_run[dir] + rank_w;
                    for (int k = 0; k < w; ++k) rank += s_wcnt[k][dir];
                    if (!FUSED) {
                        B.code[i] = (rank << 5) | dir;
                    } else if (dir == kDirStay) {
                        B.code[ps + rank] = i;
                    } else {
                        uint32_t mk = s_mrun + (uint32_t)__popcll(mm & lt_mask);
                        for (int k = 0; k < w; ++k) mk += s_mcnt[k];
                        B.mvlist[ps + mk] = (i << 5) | dir;
                    }
                }
                __syncthreads();
                if (t < kDirs) {
                    uint32_t add = 0;
#pragma unroll
                    for (int k = 0; k < NW; ++k) add += s_wcnt[k][t];
                    s_run[t] += add;
                }
                if (FUSED && t == 0) {
                    uint32_t add = 0;
#pragma unroll
                    for (int k = 0; k < NW; ++k) add += s_mcnt[k];
                    s_mrun += add;
                }
                __syncthreads();
            }
        }
        if (t < kDirs) {
            const uint32_t c = s_run[t];
            B.cnt27[(int64_t)pr.p * kDirs + t] = c;
            if (c) {
                const int64_t lD = l + (t / 9 - 1) * nb2 + ((t / 3) % 3 - 1) * nb + (t % 3 - 1);
                atomicAdd(&B.dcount[lD], c);
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    atomicOr(&B.lat_flag[lD + (e >> 2) * nb2 + ((e >> 1) & 1) * nb + (e & 1)], e == 0 ? 3u : 2u);
            }
        }
        __syncthreads();
    }
    if (bad) atomicOr(&st->domain_error, 1u);
    if (cfl) atomicOr(&st->cfl_error, 1u);
}

// First slots of the particles of old block b (lattice index l, coords lb) in the
// new stable block order, per direction t of their move: the destination block's
// first particle plus the particles arriving there from the source blocks that
// precede b in lattice order (D + e for the 26 - t offsets e before -d).  Only the
// n_act directions in which particles leave b (s_dir, from pass 1's counts) are
// needed; their (direction, source) pairs are spread over the workgroup, one
// pair per thread and round, each writing the count it finds into s_pair (26
// slots per active direction).  Indices are 32-bit (the lattice holds < 2^26
// blocks).
// With extra parts (nx != 0) a source block's count adds its extra parts' counts.
template <int NT>
__device__ __forceinline__ void move_base_pairs(int64_t l, const int lb[3], int nb, const int32_t *lat_pid,
                                                const uint8_t *lat_xf, const uint32_t *cnt27, const PartView &V,
                                                uint32_t nx, const uint32_t *s_dir, int n_act, uint32_t *s_pair) {
    const int nb2 = nb * nb, li = (int)l;
#pragma unroll 2
    for (int j = threadIdx.x; j < n_act * 26; j += NT) {
        const int k = j / 26, q = j - k * 26;
        const int t = (int)s_dir[k];
        const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
        const int ex = q / 9 - 1, ey = (q / 3) % 3 - 1, ez = q % 3 - 1;
        const int sx = lb[0] + dx + ex, sy = lb[1] + dy + ey, sz = lb[2] + dz + ez;
        const bool ok = q < 26 - t && sx >= 0 && sy >= 0 && sz >= 0 && sx < nb && sy < nb && sz < nb;
        const int idx = ok ? li + (dx + ex) * nb2 + (dy + ey) * nb + (dz + ez) : li;
        const int32_t src = lat_pid[idx];
        uint32_t c = cnt27[(int64_t)(ok && src >= 0 ? src : 0) * kDirs + (26 - q)];
        if (nx != 0u && ok && src >= 0 && lat_xf[idx]) {
            const uint32_t x0 = V.pblk_x[src], x1 = V.pblk_x[src + 1];
            for (uint32_t k = x0; k < x1; ++k) c += cnt27[(int64_t)(V.parts_cap - 1u - k) * kDirs + (26 - q)];
        }
        s_pair[j] = ok && src >= 0 ? c : 0u;
    }
}

// First particle of the destination block in direction t (0 if that block is
// outside the lattice or receives no particles; such a direction has no movers).
__device__ __forceinline__ uint32_t move_dest_start(int t, int64_t l, const int lb[3], int nb,
                                                    const int32_t *n_lat_pid, const uint32_t *n_pblk_start) {
    const int nb2 = nb * nb;
    const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
    const int X = lb[0] + dx, Y = lb[1] + dy, Z = lb[2] + dz;
    const bool in = X >= 0 && Y >= 0 && Z >= 0 && X < nb && Y < nb && Z < nb;
    const int32_t jD = n_lat_pid[in ? (int)l + dx * nb2 + dy * nb + dz : (int)l];
    const uint32_t s0 = n_pblk_start[in && jD >= 0 ? jD : 0];
    return in && jD >= 0 ? s0 : 0u;
}

// What G2P pass 2 reads and writes (only the fields it uses, to keep the
// kernel's uniform registers few).
struct UpdateArgs {
    const float *src;             // resident state (old order)
    float *dst;                   // next state (scatter) or src (in place)
    const uint32_t *counts;       // old partition: block and extra-part counts
    const uint32_t *pblk_lat, *pblk_start, *pblk_scene;
    PartView parts;
    const int32_t *lat_gid, *lat_pid;
    const uint8_t *lat_xf;
    const int32_t *n_lat_pid;     // new partition
    const uint32_t *n_pblk_start;
    const uint32_t *cnt27, *code;  // from pass 1
    const float *grid;
    Lattice L;
    const float *gstar;            // FLIP transfers: the vstar grid
    int aux;                       // kAux* channels the scatter carries
    int nonaffine;                 // FLIP without the affine term
};

// G2P update (Transfer.cu:1145-1172): v, C from the grid, x += v dt,
// F <- (I + dt C) F (J-fluid: J <- (1 + dt tr C) J).  SCATTER writes each
// particle to its slot in the next stable block order: the first slot of its
// (block, direction) group plus its stable rank in that group, both from pass 1;
// otherwise the update is in place.  A block's first 512 particles are loaded
// before its grid velocities are staged; a fixed-corotated particle (flag bit of
// the material word) then loads the rest of F.
// AFLIP: the affine FLIP update (Transfer.cu:1175-1214): v from the grid's v,
// vstar and C from its vstar, then aflip_blend with the particle's own v and
// beta.  With u.nonaffine, FLIP (Transfer.cu:1215-1250): the velocity gradient
// sum vstar_i grad W_i^T in place of C for F (C itself is kept), beta 0.
template <bool SCATTER, int NT, bool AFLIP = false>
__global__ void __launch_bounds__(NT) k_g2p_update(UpdateArgs u) {
    __shared__ float4 s_v[kTileNodes];
    __shared__ float4 s_vs[AFLIP ? kTileNodes : 1];
    __shared__ uint32_t s_base[32];
    __shared__ uint32_t s_pair[kDirs * 26];
    __shared__ uint32_t s_dir[kDirs];  // the block's move directions with particles, in order
    __shared__ uint32_t s_nact;
    const int t = threadIdx.x;
    const uint32_t nblocks = u.counts[0], nx = u.counts[2];
    const uint32_t nunits = nblocks + nx;
    for (uint32_t j = xcd_first(nunits); j < nunits; j += gridDim.x) {  // parts in XCD slab order
        const PartRef pr = part_ref(u.parts, j, nblocks);
        const uint32_t b = pr.b;
        const SceneParams sc = scene_params(u.L, u.L.nscenes == 1 ? 0u : u.pblk_scene[b]);
        const int64_t l = u.pblk_lat[b];
        const uint32_t ps = pr.lo, pe = pr.hi;
        // first two chunks: code, x, material word, F00|J, caller index
        uint32_t qcode[2], qmw[2], qpid[2];
        float qx[2][3], qf0[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const uint32_t i = min(ps + q * NT + t, pe - 1);
            qcode[q] = SCATTER ? u.code[i] : 0u;
#pragma unroll
            for (int k = 0; k < 3; ++k) qx[q][k] = chan(u.src, i, kChX + k);
            qmw[q] = chan_u(u.src, i, kChMat);
            qf0[q] = chan(u.src, i, kChF);
            qpid[q] = SCATTER ? chan_u(u.src, i, kChPid) : 0u;
        }
        int lb[3];
        lattice_coords(sc, l, lb[0], lb[1], lb[2]);
        const int bc[3] = {(lb[0] + u.L.origin) * 4, (lb[1] + u.L.origin) * 4, (lb[2] + u.L.origin) * 4};
        const int64_t nb = sc.nb, nb2 = nb * nb;
        if (SCATTER && t < 64) {
            // the directions in which this block's particles leave (pass 1's counts)
            const bool act = t < kDirs && u.cnt27[(int64_t)pr.p * kDirs + t] != 0u;
            const uint64_t mask = __ballot(act);
            if (act) s_dir[__popcll(mask & ((1ull << t) - 1ull))] = (uint32_t)t;
            if (t == 0) s_nact = (uint32_t)__popcll(mask);
        }
        stage_block_velocity<NT>(s_v, u.grid, u.lat_gid, l, nb, nb2);
        if (AFLIP) stage_block_velocity<NT>(s_vs, u.gstar, u.lat_gid, l, nb, nb2);
        __syncthreads();
        if (SCATTER) {
            const int n_act = (int)s_nact;
            move_base_pairs<NT>(l, lb, sc.nb, u.lat_pid, u.lat_xf, u.cnt27, u.parts, nx, s_dir, n_act, s_pair);
            const uint32_t dest0 = t < n_act ? move_dest_start((int)s_dir[t], l, lb, sc.nb, u.n_lat_pid,
                                                                 u.n_pblk_start)
                                             : 0u;
            __syncthreads();
            if (t < n_act) {
                uint32_t base = dest0;
#pragma unroll
                for (int q = 0; q < 26; ++q) base += s_pair[t * 26 + q];
                if (pr.k != 0xffffffffu) {  // extra part q of its block: after the first part and extras 1 .. q-1
                    base += u.cnt27[(int64_t)b * kDirs + s_dir[t]];
                    const uint32_t q = (pr.lo - u.pblk_start[b]) >> u.parts.shift;
                    for (uint32_t k = pr.k + 1u - q; k < pr.k; ++k)
                        base += u.cnt27[(int64_t)(u.parts.parts_cap - 1u - k) * kDirs + s_dir[t]];
                }
                s_base[s_dir[t]] = base;
            }
            __syncthreads();
        }
        for (uint32_t c0 = ps, q = 0; c0 < pe; c0 += NT, ++q) {
            const uint32_t i = c0 + t;
            if (i >= pe) break;
            uint32_t code, mw, pid;
            float xp[3], f0;
            if (q < 2) {
                code = q == 0 ? qcode[0] : qcode[1];
                mw = q == 0 ? qmw[0] : qmw[1];
                pid = q == 0 ? qpid[0] : qpid[1];
                f0 = q == 0 ? qf0[0] : qf0[1];
#pragma unroll
                for (int k = 0; k < 3; ++k) xp[k] = q == 0 ? qx[0][k] : qx[1][k];
            } else {
                code = SCATTER ? u.code[i] : 0u;
#pragma unroll
                for (int k = 0; k < 3; ++k) xp[k] = chan(u.src, i, kChX + k);
                mw = chan_u(u.src, i, kChMat);
                f0 = chan(u.src, i, kChF);
                pid = SCATTER ? chan_u(u.src, i, kChPid) : 0u;
            }
            const uint32_t dir = code & 31u;
            const uint32_t dst = SCATTER ? s_base[dir] + (code >> 5) : i;
            const bool fcr = (mw & kMatF) != 0u;
            float Fo[3][3];
            if (fcr) {
                Fo[0][0] = f0;
#pragma unroll
                for (int q9 = 1; q9 < 9; ++q9) Fo[q9 / 3][q9 % 3] = chan(u.src, i, kChF + q9);
            }
            Arena ar;
            make_arena(xp, sc.dxinv, ar);
            const int lc[3] = {clamp03(ar.base[0] - bc[0]), clamp03(ar.base[1] - bc[1]), clamp03(ar.base[2] - bc[2])};
            float vel[3], Cn[3][3], xn[3];
            if (AFLIP) {
                float v0[3], vg[3], vst[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) v0[k] = chan(u.src, i, kChV + k);
                g2p_gather<false>(s_v, ar, lc, sc.dx, sc.Dinv, vg, Cn);
                g2p_gather<true>(s_vs, ar, lc, sc.dx, u.nonaffine ? sc.dxinv : sc.Dinv, vst, Cn, u.nonaffine != 0);
                aflip_blend(xp, v0, vg, vst, u.nonaffine ? 0.0f : chan(u.src, i, kChBeta), sc.dt, xn, vel);
            } else {
                g2p_gather<true>(s_v, ar, lc, sc.dx, sc.Dinv, vel, Cn);
#pragma unroll
                for (int k = 0; k < 3; ++k) xn[k] = xp[k] + vel[k] * sc.dt;
            }
            if (fcr) {
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        // F <- (I + dt C) F (Transfer.cu:1169-1172)
                        const float g0 = (r == 0 ? 1.0f : 0.0f) + Cn[r][0] * sc.dt;
                        const float g1 = (r == 1 ? 1.0f : 0.0f) + Cn[r][1] * sc.dt;
                        const float g2 = (r == 2 ? 1.0f : 0.0f) + Cn[r][2] * sc.dt;
                        chan(u.dst, dst, kChF + 3 * r + c) = g0 * Fo[0][c] + g1 * Fo[1][c] + g2 * Fo[2][c];
                    }
            } else {
                chan(u.dst, dst, kChF) = (1.0f + sc.dt * (Cn[0][0] + Cn[1][1] + Cn[2][2])) * f0;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                chan(u.dst, dst, kChX + k) = xn[k];
                chan(u.dst, dst, kChV + k) = vel[k];
            }
            if (!AFLIP || !u.nonaffine) {
#pragma unroll
                for (int q9 = 0; q9 < 9; ++q9) chan(u.dst, dst, kChC + q9) = Cn[q9 / 3][q9 % 3];
            } else if (SCATTER) {
#pragma unroll
                for (int q9 = 0; q9 < 9; ++q9) chan(u.dst, dst, kChC + q9) = chan(u.src, i, kChC + q9);
            }
            if (SCATTER) {
                set_u(u.dst, dst, kChMat, mw);
                set_u(u.dst, dst, kChPid, pid);
                if (u.aux & kAuxBeta) chan(u.dst, dst, kChBeta) = chan(u.src, i, kChBeta);
                if (u.aux & kAuxLogJp) chan(u.dst, dst, kChLogJp) = chan(u.src, i, kChLogJp);
                if (u.aux & kAuxLame) {
                    chan(u.dst, dst, kChMu) = chan(u.src, i, kChMu);
                    chan(u.dst, dst, kChLam) = chan(u.src, i, kChLam);
                }
                if (u.aux & kAuxFiber) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) chan(u.dst, dst, kChFiber + k) = chan(u.src, i, kChFiber + k);
                }
                if (u.aux & kAuxMass) {
                    chan(u.dst, dst, kChMass) = chan(u.src, i, kChMass);
                    chan(u.dst, dst, kChVol) = chan(u.src, i, kChVol);
                }
            }
        }
        __syncthreads();
    }
}

// What the fused G2P / P2G pass reads and writes.
struct FusedArgs {
    const float *src;             // resident state (old order)
    float *dst;                   // next state (new order)
    const uint32_t *pblk_start;   // old partition
    const int32_t *lat_pid, *lat_gid;
    const uint32_t *n_counts;     // new partition (from the lattice scan)
    const uint32_t *n_pblk_lat, *n_pblk_start, *n_pblk_scene;
    const uint32_t *cnt27, *stay, *mv;  // the move pass's counts, stayer and mover lists
    uint32_t *arr;                // scratch: old index of each arriving particle, at its new slot
    const float *grid;            // this substep's updated grid velocities
    float *tiles;                 // the next substep's P2G tiles (new blocks)
    const DevMaterial *mats;
    Lattice L;
    DevStatus *st;
    int aux;                      // kAux* channels carried
    int write_vc;                 // store v and C too (the last substep of a call)
    uint32_t n;                   // resident particles
};

// The fused pass: G2P of this substep (Transfer.cu:1145-1172) and the next
// substep's stress and P2G (Transfer.cu:844-890), one workgroup per block of
// the NEW partition.  A new block's particles are, in the new stable order,
// the movers of the source blocks before it in lattice order, its own stayers,
// then the movers of the later sources; each source contributes its movers of
// the matching direction in its own order.  The workgroup lists the arrivals'
// old indices (from the sources' mover lists) at their new slots, stages the
// updated grid velocities around the block for its stayers (arrivals gather
// theirs from HBM, the same arithmetic), and walks the new slots in chunks of
// kP2GChunk: per particle the G2P update (x, v, C, F or J; the return mapping of
// plastic particles), the new state written to its slot (v and C only when
// write_vc: nothing else reads them before the next G2P overwrites them), then
// from the registers the stress term and the P2G payload, staged in LDS with
// the particle's cell and stable in-cell rank; the chunk then goes through the
// P2G phases (p2g_chunk_tail, through the cell-order index s_sorted).  Chunks
// are the same new-slot ranges as k_p2g's over the new state, and every
// floating-point operation is the one the unfused kernels do, so the new
// state and the tiles equal those of G2P update + k_stress + k_p2g bit for bit.
// GENERAL: every model, plasticity and the ARAP fiber term; otherwise the
// context's solids are all fixed corotated without either, and the FP64 code of
// the others (and its registers) is left out.
// amdgpu_waves_per_eu(3): at most 168 VGPRs, so four workgroups fit on a CU.
template <bool PP, bool GENERAL>
__global__ void __launch_bounds__(kP2GThreads) __attribute__((amdgpu_waves_per_eu(3))) k_g2p2g(FusedArgs u) {
    __shared__ float s_pl[kPL * kP2GChunk];
    __shared__ float4 s_v[kTileNodes];
    __shared__ uint32_t s_wcnt[kP2GRounds * 3][64];
    __shared__ uint32_t s_cstart[65];
    __shared__ uint8_t s_task[3 * 64];
    __shared__ uint16_t s_sorted[kP2GChunk];
    __shared__ uint32_t s_rb[kDirs + 1];  // first new slot (relative) of each source's run, and the total
    __shared__ int32_t s_src[kDirs];      // old block of each source (-1: none)

    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t nblocks = u.n_counts[0];
    bool mismatch = false;
    s_task[t] = (uint8_t)(t / 3);
    for (uint32_t j = xcd_first(nblocks); j < nblocks; j += gridDim.x) {  // blocks in XCD slab order
        const int64_t l = u.n_pblk_lat[j];
        const SceneParams sc = scene_params(u.L, u.L.nscenes == 1 ? 0u : u.n_pblk_scene[j]);
        const int nb = sc.nb, nb2 = nb * nb;
        int lb[3];
        lattice_coords(sc, l, lb[0], lb[1], lb[2]);
        const int bc[3] = {(lb[0] + u.L.origin) * 4, (lb[1] + u.L.origin) * 4, (lb[2] + u.L.origin) * 4};
        const uint32_t ns = u.n_pblk_start[j], nD = u.n_pblk_start[j + 1] - ns;
        // the 27 sources D + e in lattice order (q = 9 (ex + 1) + 3 (ey + 1) + ez + 1), moving in direction 26 - q
        if (w == 0) {
            uint32_t c = 0u;
            int32_t src = -1;
            if (lane < kDirs) {
                const int ex = lane / 9 - 1, ey = (lane / 3) % 3 - 1, ez = lane % 3 - 1;
                const int sx = lb[0] + ex, sy = lb[1] + ey, sz = lb[2] + ez;
                if (sx >= 0 && sy >= 0 && sz >= 0 && sx < nb && sy < nb && sz < nb)
                    src = u.lat_pid[(int)l + ex * nb2 + ey * nb + ez];
                c = src >= 0 ? u.cnt27[(int64_t)src * kDirs + (26 - lane)] : 0u;
            }
            uint32_t inc = c;
#pragma unroll
            for (int off = 1; off < 32; off <<= 1) {
                const uint32_t y = __shfl_up(inc, off, 64);
                if (lane >= off) inc += y;
            }
            if (lane < kDirs) {
                s_rb[lane] = inc - c;
                s_src[lane] = src;
            }
            if (lane == kDirs - 1) s_rb[kDirs] = inc;
        }
        __syncthreads();
        const uint32_t n_pre = s_rb[13], n_stay = s_rb[14] - s_rb[13];
        mismatch |= s_rb[kDirs] != nD;  // must never happen
        const int32_t sD = s_src[13];   // the old block at this lattice entry (-1: none)
        const uint32_t psD = sD >= 0 ? u.pblk_start[sD] : 0u;
        // arrivals: per source, its movers in this direction, in its order, at their new slots
        for (int q = w; q < kDirs; q += 3) {
            const uint32_t c = s_rb[q + 1] - s_rb[q];
            if (q == 13 || c == 0u) continue;
            const int32_t S = s_src[q];
            const uint32_t pS = u.pblk_start[S];
            const uint32_t nmov = (u.pblk_start[S + 1] - pS) - u.cnt27[(int64_t)S * kDirs + 13];
            const uint32_t want = (uint32_t)(26 - q), base = ns + s_rb[q];
            uint32_t got = 0u;
            for (uint32_t k0 = 0; k0 < nmov && got < c; k0 += 64) {
                const uint32_t k = k0 + (uint32_t)lane;
                const uint32_t e = k < nmov ? u.mv[pS + k] : 31u;
                const bool hit = (e & 31u) == want;
                const uint64_t hm = __ballot(hit);
                const uint32_t slot = got + count_below(hm);
                if (hit && slot < c) u.arr[base + slot] = e >> 5;  // (more matches than counted: never)
                got += (uint32_t)__popcll(hm);
            }
        }
        if (n_stay > 0u) stage_block_velocity<kP2GThreads>(s_v, u.grid, u.lat_gid, l, nb, nb2);
        __syncthreads();  // also orders the arrival list's global stores before its loads below
        float *tb = u.tiles + (int64_t)j * (4 * kTileNodes);
        const float dx = sc.dx, dxinv = sc.dxinv, dt = sc.dt;
        const float sdt = -dt * sc.Dinv;
        for (uint32_t c0 = 0; c0 < nD; c0 += kP2GChunk) {
            int cell[kP2GRounds];
            uint32_t rank[kP2GRounds];
#pragma unroll
            for (int r = 0; r < kP2GRounds; ++r) {
                const uint32_t k = c0 + r * kP2GThreads + t;
                const bool valid = k < nD;
                cell[r] = 0;
                if (valid) {
                    const bool stayer = k - n_pre < n_stay;  // unsigned: k in [n_pre, n_pre + n_stay)
                    uint32_t i = stayer ? u.stay[psD + (k - n_pre)] : u.arr[ns + k];
                    if (i >= u.n) {  // inconsistent lists: never; kept inside the buffers
                        mismatch = true;
                        i = 0;
                    }
                    const uint32_t dst = ns + k;
                    float xp[3];
#pragma unroll
                    for (int q = 0; q < 3; ++q) xp[q] = chan(u.src, i, kChX + q);
                    const uint32_t mw = chan_u(u.src, i, kChMat);
                    const bool fcr = (mw & kMatF) != 0u;
                    // ---- G2P
                    Arena ar;
                    make_arena(xp, dxinv, ar);
                    float vel[3], Cn[3][3];
                    if (stayer) {
                        const int lc[3] = {clamp03(ar.base[0] - bc[0]), clamp03(ar.base[1] - bc[1]),
                                           clamp03(ar.base[2] - bc[2])};
                        g2p_gather<true, true>(s_v, ar, lc, dx, sc.Dinv, vel, Cn);
                    } else {
                        g2p_gather_fn<true, true>(
                            [&](int ox, int oy, int oz) {
                                return grid_node_velocity(u.grid, u.lat_gid, sc, u.L.origin, ar.base[0] + ox,
                                                          ar.base[1] + oy, ar.base[2] + oz);
                            },
                            ar, dx, sc.Dinv, vel, Cn);
                    }
                    float *pl = s_pl + r * kP2GThreads + t;
                    // the advected position, its cell in this block and fractional position
                    float xn[3];
#pragma unroll
                    for (int q = 0; q < 3; ++q) xn[q] = xp[q] + vel[q] * dt;
                    {
                        int lc[3];
#pragma unroll
                        for (int q = 0; q < 3; ++q) {
                            chan(u.dst, dst, kChX + q) = xn[q];
                            const float X = xn[q] * dxinv;
                            const int base = cell_floor(X);
                            pl[(1 + q) * kP2GChunk] = X - (float)base;
                            lc[q] = base - bc[q];
                            mismatch |= (lc[q] < 0 || lc[q] > 3);
                            lc[q] = clamp03(lc[q]);
                        }
                        cell[r] = (lc[0] * 4 + lc[1]) * 4 + lc[2];
                    }
                    float Fn[3][3], Jn = 0.0f;
                    if (fcr) {
                        float Fo[3][3];
#pragma unroll
                        for (int q9 = 0; q9 < 9; ++q9) Fo[q9 / 3][q9 % 3] = chan(u.src, i, kChF + q9);
#pragma unroll
                        for (int r2 = 0; r2 < 3; ++r2)
#pragma unroll
                            for (int c = 0; c < 3; ++c) {
                                // F <- (I + dt C) F (Transfer.cu:1169-1172)
                                const float g0 = (r2 == 0 ? 1.0f : 0.0f) + Cn[r2][0] * dt;
                                const float g1 = (r2 == 1 ? 1.0f : 0.0f) + Cn[r2][1] * dt;
                                const float g2 = (r2 == 2 ? 1.0f : 0.0f) + Cn[r2][2] * dt;
                                Fn[r2][c] = g0 * Fo[0][c] + g1 * Fo[1][c] + g2 * Fo[2][c];
                            }
                    } else {
                        Jn = (1.0f + dt * (Cn[0][0] + Cn[1][1] + Cn[2][2])) * chan(u.src, i, kChF);
                    }
                    // v and C wait in the payload's slots (4..15) while the stress is formed
#pragma unroll
                    for (int q = 0; q < 3; ++q) pl[(4 + q) * kP2GChunk] = vel[q];
#pragma unroll
                    for (int q9 = 0; q9 < 9; ++q9) pl[(7 + q9) * kP2GChunk] = Cn[q9 / 3][q9 % 3];
                    if (u.write_vc) {
#pragma unroll
                        for (int q = 0; q < 3; ++q) chan(u.dst, dst, kChV + q) = vel[q];
#pragma unroll
                        for (int q9 = 0; q9 < 9; ++q9) chan(u.dst, dst, kChC + q9) = Cn[q9 / 3][q9 % 3];
                    }
                    const DevMaterial mt = u.mats[mw & kMatIndexMask];
                    // ZSReturnMapping after G2P (MPMPipeline.cu:953-972)
                    float logJp = (u.aux & kAuxLogJp) ? chan(u.src, i, kChLogJp) : 0.0f;
                    if (GENERAL && (mw & kMatPlastic)) plastic_project(mt, Fn, logJp);
                    // ---- the rest of the new state at its slot
                    if (fcr) {
#pragma unroll
                        for (int q9 = 0; q9 < 9; ++q9) chan(u.dst, dst, kChF + q9) = Fn[q9 / 3][q9 % 3];
                    } else {
                        chan(u.dst, dst, kChF) = Jn;
                    }
                    set_u(u.dst, dst, kChMat, mw);
                    set_u(u.dst, dst, kChPid, chan_u(u.src, i, kChPid));
                    if (u.aux & kAuxBeta) chan(u.dst, dst, kChBeta) = chan(u.src, i, kChBeta);
                    if (u.aux & kAuxLogJp) chan(u.dst, dst, kChLogJp) = logJp;
                    if (u.aux & kAuxLame) {
                        chan(u.dst, dst, kChMu) = chan(u.src, i, kChMu);
                        chan(u.dst, dst, kChLam) = chan(u.src, i, kChLam);
                    }
                    float fa[3] = {0.0f, 0.0f, 0.0f};
                    if (u.aux & kAuxFiber) {
#pragma unroll
                        for (int q = 0; q < 3; ++q) {
                            fa[q] = chan(u.src, i, kChFiber + q);
                            chan(u.dst, dst, kChFiber + q) = fa[q];
                        }
                    }
                    if (u.aux & kAuxMass) {
                        chan(u.dst, dst, kChMass) = chan(u.src, i, kChMass);
                        chan(u.dst, dst, kChVol) = chan(u.src, i, kChVol);
                    }
                    // ---- the next substep's stress term and P2G payload
                    float sg[3][3];
                    if (fcr) {
                        const float s = sdt * (PP ? chan(u.src, i, kChVol) : mt.vol);
                        const float mu = u.L.pp_lame ? chan(u.src, i, kChMu) : mt.mu;
                        const float lam = u.L.pp_lame ? chan(u.src, i, kChLam) : mt.lam;
                        if (GENERAL) {
                            f_stress_term(mt, s, mu, lam, Fn, fa, sg);
                        } else {  // f_stress_term's fixed-corotated branch
                            float Pk[3][3];
                            fcr_piola(Fn, mu, lam, Pk);
#pragma unroll
                            for (int r2 = 0; r2 < 3; ++r2)
#pragma unroll
                                for (int c = 0; c < 3; ++c)
                                    sg[r2][c] = (s * Pk[r2][0]) * Fn[c][0] + (s * Pk[r2][1]) * Fn[c][1] +
                                                (s * Pk[r2][2]) * Fn[c][2];
                        }
                    }
                    P2GIn in;
                    float fr[3];
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        fr[q] = pl[(1 + q) * kP2GChunk];
                        in.v[q] = pl[(4 + q) * kP2GChunk];
                    }
#pragma unroll
                    for (int q9 = 0; q9 < 9; ++q9) in.C[q9] = pl[(7 + q9) * kP2GChunk];
                    in.J = Jn;
                    in.mw = mw;
                    in.pp = PP ? u.src : nullptr;
                    p2g_payload_f<kP2GChunk, 0>(in, i, u.mats, [&](int q) { return sg[q / 3][q % 3]; }, dx, sdt,
                                                fr, pl);
                }
                // stable rank among the chunk's particles of the same cell (6 wave ballots)
                const uint64_t match = same_key_lanes<6>((uint32_t)cell[r], valid);
                rank[r] = valid ? count_below(match) : 0xffffffffu;
                s_wcnt[r * 3 + w][lane] = 0u;
                wave_lds_order();
                if (valid && rank[r] == 0) s_wcnt[r * 3 + w][cell[r]] = (uint32_t)__popcll(match);
            }
            __syncthreads();
            if (w == 0) {
                uint32_t tot = 0;
#pragma unroll
                for (int k = 0; k < kP2GRounds * 3; ++k) tot += s_wcnt[k][lane];
                uint32_t inc = tot;
#pragma unroll
                for (int off = 1; off < 64; off <<= 1) {
                    const uint32_t y = __shfl_up(inc, off, 64);
                    if (lane >= off) inc += y;
                }
                s_cstart[lane] = inc - tot;
                if (lane == 63) s_cstart[64] = inc;
                if (c0 == 0) {
                    // phase B's tasks sorted by the first chunk's cell counts, as k_p2g does
                    const uint32_t key = min(tot, 1023u);
                    uint64_t eq = ~0ull;
                    uint32_t gt = 0;
#pragma unroll
                    for (int bit = 9; bit >= 0; --bit) {
                        const bool on = (key >> bit) & 1;
                        const uint64_t bb = __ballot(on);
                        if (!on) gt += (uint32_t)__popcll(eq & bb);
                        eq &= on ? bb : ~bb;
                    }
                    const uint32_t rk = gt + count_below(eq);
#pragma unroll
                    for (int k = 0; k < 3; ++k) s_task[3 * rk + k] = (uint8_t)lane;
                }
            }
            __syncthreads();
            // cell-order position -> the payload's slot
#pragma unroll
            for (int r = 0; r < kP2GRounds; ++r) {
                if (rank[r] == 0xffffffffu) continue;
                uint32_t pos = s_cstart[cell[r]] + rank[r];
                for (int k = 0; k < r * 3 + w; ++k) pos += s_wcnt[k][cell[r]];
                s_sorted[pos] = (uint16_t)(r * kP2GThreads + t);
            }
            __syncthreads();
            p2g_chunk_tail<kP2GForward>(s_pl, s_task, s_cstart, s_sorted, tb, c0 == 0);
        }
    }
    if (mismatch) atomicOr(&u.st->bin_mismatch, 1u);
}

// ------------------------------------------------------------------ import / export
__device__ __forceinline__ int64_t view_index(int layout, int64_t n, int tile_ch, int chn, int nch, int c, int64_t i) {
    if (layout == ZMPM_LAYOUT_SOA) return (int64_t)c * n + i;
    if (layout == ZMPM_LAYOUT_AOS) return i * nch + c;
    return ((i >> 5) * tile_ch + chn + c) * 32 + (i & 31);
}

// Particle j of the view becomes resident particle off + j, caller index off + j.
__global__ void __launch_bounds__(kThreads) k_import(State dst, int64_t cap, zmpm_particles v, int64_t off,
                                                     const DevMaterial *mats, int32_t nmat, DevStatus *st) {
    const int64_t n = v.n;
    bool badmat = false;
    for (int64_t j = blockIdx.x * (int64_t)kThreads + threadIdx.x; j < n; j += (int64_t)gridDim.x * kThreads) {
        const int64_t i = off + j;
        int32_t mid = v.material ? v.material[j] : v.default_material;
        if (mid < 0 || mid >= nmat) {
            badmat = true;
            mid = 0;
        }
        const int model = mats[mid].model;
        const bool fbased = model != ZMPM_MODEL_JFLUID;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            chan(dst.ch, i, c) = v.x[view_index(v.layout, n, v.tile_channels, v.x_chn, 3, c, j)];
            chan(dst.ch, i, 3 + c) = v.v ? v.v[view_index(v.layout, n, v.tile_channels, v.v_chn, 3, c, j)] : 0.0f;
        }
#pragma unroll
        for (int c = 0; c < 9; ++c)
            chan(dst.ch, i, 6 + c) = v.C ? v.C[view_index(v.layout, n, v.tile_channels, v.C_chn, 9, c, j)] : 0.0f;
        if (fbased) {
#pragma unroll
            for (int c = 0; c < 9; ++c)
                chan(dst.ch, i, 15 + c) =
                    v.F ? v.F[view_index(v.layout, n, v.tile_channels, v.F_chn, 9, c, j)] : ((c % 4 == 0) ? 1.0f : 0.0f);
        } else {
            chan(dst.ch, i, 15) = v.J ? v.J[view_index(v.layout, n, v.tile_channels, v.J_chn, 1, 0, j)] : 1.0f;
#pragma unroll
            for (int c = 1; c < 9; ++c) chan(dst.ch, i, 15 + c) = 0.0f;
        }
        set_u(dst.ch, (uint32_t)i, kChMat,
              (uint32_t)mid | (fbased ? kMatF : 0u) | (mats[mid].plasticity != ZMPM_PLASTIC_NONE ? kMatPlastic : 0u));
        set_u(dst.ch, (uint32_t)i, kChPid, (uint32_t)i);
        chan(dst.ch, i, kChBeta) =
            v.beta ? v.beta[view_index(v.layout, n, v.tile_channels, v.beta_chn, 1, 0, j)] : mats[mid].flip_beta;
        chan(dst.ch, i, kChLogJp) =
            v.logJp ? v.logJp[view_index(v.layout, n, v.tile_channels, v.logJp_chn, 1, 0, j)] : kLogJpInit;
        const bool pp = v.mu && v.lam;
        chan(dst.ch, i, kChMu) = pp ? v.mu[view_index(v.layout, n, v.tile_channels, v.mu_chn, 1, 0, j)] : mats[mid].mu;
        chan(dst.ch, i, kChLam) =
            pp ? v.lam[view_index(v.layout, n, v.tile_channels, v.lam_chn, 1, 0, j)] : mats[mid].lam;
        // the fiber "a", (0, 1, 0) when none is given (Generation.cpp:982)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            chan(dst.ch, i, kChFiber + c) =
                v.a ? v.a[view_index(v.layout, n, v.tile_channels, v.a_chn, 3, c, j)] : (c == 1 ? 1.0f : 0.0f);
        // per-particle mass and volume (Transfer.cu:857-858), else the material's
        const bool pm = v.m && v.vol;
        chan(dst.ch, i, kChMass) = pm ? v.m[view_index(v.layout, n, v.tile_channels, v.m_chn, 1, 0, j)] : mats[mid].mass;
        chan(dst.ch, i, kChVol) = pm ? v.vol[view_index(v.layout, n, v.tile_channels, v.vol_chn, 1, 0, j)] : mats[mid].vol;
    }
    if (badmat) atomicOr(&st->bad_material, 1u);
}

// ------------------------------------------------------------------ Morton order
// Every third bit from the 10 low bits of v (the multiply-and-mask spreading of
// Karras 2012, "Maximizing parallelism in the construction of BVHs, octrees, and
// k-d trees", section 4).
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// Channel c of particle i of a whole buffer of nch channels: SoA planes, AoS records,
// or the AoSoA32 TileVector (view_index addresses one attribute's array instead).
__device__ __forceinline__ int64_t buf_index(int layout, int64_t n, int nch, int c, int64_t i) {
    if (layout == ZMPM_LAYOUT_SOA) return (int64_t)c * n + i;
    if (layout == ZMPM_LAYOUT_AOS) return i * nch + c;
    return ((i >> 5) * nch + c) * 32 + (i & 31);
}

// ZenoParticles::orderByMortonCode's codes (projects/CUDA/Structures.hpp:423-427): the
// position (channels x_chn..x_chn+2 of the view) in the box's uniform coordinates
// (x - lo) / ext (AABBBox::getUniformCoord, correctly rounded), per axis
// trunc(min(max(c * 1024, 0), 1023)), interleaved x y z from the high bit.
__global__ void __launch_bounds__(kThreads) k_morton_codes(int32_t layout, int32_t nch, int64_t n, const float *src,
                                                           int32_t x_chn, float lx, float ly, float lz, float ex,
                                                           float ey, float ez, uint32_t *codes) {
    const float lo[3] = {lx, ly, lz}, ext[3] = {ex, ey, ez};
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        uint32_t q[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float x = src[buf_index(layout, n, nch, x_chn + k, i)];
            const float c = __fdiv_rn(x - lo[k], ext[k]);
            q[k] = (uint32_t)fminf(fmaxf(c * 1024.0f, 0.0f), 1023.0f);
        }
        codes[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    }
}

// Every channel of particle order[k] of src becomes particle k of dst (Structures.hpp:438-443).
__global__ void __launch_bounds__(kThreads) k_permute_channels(int32_t layout, int32_t nch, int64_t n,
                                                               const float *src, float *dst, const uint32_t *order) {
    for (int64_t k = blockIdx.x * (int64_t)kThreads + threadIdx.x; k < n; k += (int64_t)gridDim.x * kThreads) {
        const int64_t i = order[k];
        for (int c = 0; c < nch; ++c)
            dst[buf_index(layout, n, nch, c, k)] = src[buf_index(layout, n, nch, c, i)];
    }
}

// aux: the kAux* channels the resident state keeps; the others are exported
// as their initial values (the material's flip_beta, logJp -0.04).
// Resident particles of caller index off .. off + v.n - 1 go to the view's entries 0 .. v.n - 1.
__global__ void __launch_bounds__(kThreads) k_export(State src, int64_t n_res, zmpm_particles v, int64_t off,
                                                     const DevMaterial *mats, int aux) {
    const int64_t n = v.n;
    for (int64_t k = blockIdx.x * (int64_t)kThreads + threadIdx.x; k < n_res; k += (int64_t)gridDim.x * kThreads) {
        const int64_t i = (int64_t)chan_u(src.ch, (uint32_t)k, kChPid) - off;
        if (i < 0 || i >= n) continue;
        const uint32_t mw = chan_u(src.ch, (uint32_t)k, kChMat);
        const bool fcr = (mw & kMatF) != 0u;
        if (v.beta)
            v.beta[view_index(v.layout, n, v.tile_channels, v.beta_chn, 1, 0, i)] =
                (aux & kAuxBeta) ? chan(src.ch, k, kChBeta) : mats[mw & kMatIndexMask].flip_beta;
        if (v.logJp)
            v.logJp[view_index(v.layout, n, v.tile_channels, v.logJp_chn, 1, 0, i)] =
                (aux & kAuxLogJp) ? chan(src.ch, k, kChLogJp) : kLogJpInit;
        if (v.mu)
            v.mu[view_index(v.layout, n, v.tile_channels, v.mu_chn, 1, 0, i)] =
                (aux & kAuxLame) ? chan(src.ch, k, kChMu) : mats[mw & kMatIndexMask].mu;
        if (v.lam)
            v.lam[view_index(v.layout, n, v.tile_channels, v.lam_chn, 1, 0, i)] =
                (aux & kAuxLame) ? chan(src.ch, k, kChLam) : mats[mw & kMatIndexMask].lam;
        if (v.a) {
#pragma unroll
            for (int c = 0; c < 3; ++c)
                v.a[view_index(v.layout, n, v.tile_channels, v.a_chn, 3, c, i)] =
                    (aux & kAuxFiber) ? chan(src.ch, k, kChFiber + c) : (c == 1 ? 1.0f : 0.0f);
        }
        if (v.m)
            v.m[view_index(v.layout, n, v.tile_channels, v.m_chn, 1, 0, i)] =
                (aux & kAuxMass) ? chan(src.ch, k, kChMass) : mats[mw & kMatIndexMask].mass;
        if (v.vol)
            v.vol[view_index(v.layout, n, v.tile_channels, v.vol_chn, 1, 0, i)] =
                (aux & kAuxMass) ? chan(src.ch, k, kChVol) : mats[mw & kMatIndexMask].vol;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            v.x[view_index(v.layout, n, v.tile_channels, v.x_chn, 3, c, i)] = chan(src.ch, k, c);
            if (v.v) v.v[view_index(v.layout, n, v.tile_channels, v.v_chn, 3, c, i)] = chan(src.ch, k, 3 + c);
        }
        if (v.C) {
#pragma unroll
            for (int c = 0; c < 9; ++c)
                v.C[view_index(v.layout, n, v.tile_channels, v.C_chn, 9, c, i)] = chan(src.ch, k, 6 + c);
        }
        if (fcr) {
            if (v.F) {
#pragma unroll
                for (int c = 0; c < 9; ++c)
                    v.F[view_index(v.layout, n, v.tile_channels, v.F_chn, 9, c, i)] = chan(src.ch, k, 15 + c);
            }
        } else if (v.J) {
            v.J[view_index(v.layout, n, v.tile_channels, v.J_chn, 1, 0, i)] = chan(src.ch, k, 15);
        }
    }
}

// ------------------------------------------------------------------ return mapping
// ZSReturnMapping (MPMPipeline.cu:953-972) on the particles flagged plastic:
// the projected deformation gradient replaces F (FP64 left stretch; see
// project_hencky_strain).
__global__ void __launch_bounds__(kThreads) k_return_mapping(State A, int64_t n, const DevMaterial *mats) {
    for (int64_t k = blockIdx.x * (int64_t)kThreads + threadIdx.x; k < n; k += (int64_t)gridDim.x * kThreads) {
        const uint32_t i = (uint32_t)k;
        const uint32_t mw = chan_u(A.ch, i, kChMat);
        if (!(mw & kMatPlastic)) continue;
        const DevMaterial mt = mats[mw & kMatIndexMask];
        float F[3][3];
#pragma unroll
        for (int q = 0; q < 9; ++q) F[q / 3][q % 3] = chan(A.ch, i, kChF + q);
        const bool nacc = mt.plasticity == ZMPM_PLASTIC_CAM_CLAY;
        float logJp = nacc ? chan(A.ch, i, kChLogJp) : 0.0f;
        plastic_project(mt, F, logJp);
        // unchanged values are written back as they were read
#pragma unroll
        for (int q = 0; q < 9; ++q) chan(A.ch, i, kChF + q) = F[q / 3][q % 3];
        if (nacc) chan(A.ch, i, kChLogJp) = logJp;
    }
}

// The channels of mask (kAux*) reset from the materials: beta = flip_beta,
// logJp = the reference's initial -0.04.
__global__ void __launch_bounds__(kThreads) k_init_aux(State A, int64_t n, const DevMaterial *mats, int mask) {
    for (int64_t k = blockIdx.x * (int64_t)kThreads + threadIdx.x; k < n; k += (int64_t)gridDim.x * kThreads) {
        const uint32_t i = (uint32_t)k;
        if (mask & kAuxBeta) chan(A.ch, i, kChBeta) = mats[chan_u(A.ch, i, kChMat) & kMatIndexMask].flip_beta;
        if (mask & kAuxLogJp) chan(A.ch, i, kChLogJp) = kLogJpInit;
        if (mask & kAuxFiber) {
            chan(A.ch, i, kChFiber) = 0.0f;
            chan(A.ch, i, kChFiber + 1) = 1.0f;
            chan(A.ch, i, kChFiber + 2) = 0.0f;
        }
    }
}

// ComputeParticleBeta at x + dt v: with a level-set boundary (Boundary.cu:284-313)
// a particle inside a collider of its scene (phi < 0) and moving into it,
// (v - dbdt) . (R grad phi) < 0, is moving in (the level set carries no material
// velocity, so the object's is its translation rate); with a boundary grid
// (Boundary.cu:173-219, the boundary particles' sums bp.acc on the partition P)
// a particle whose linear stencil (base floor(x / dx), 8 nodes) reaches a node of
// boundary mass with (v - v_b) . n_b < 0 is moving in.  Moving in gives beta 0;
// otherwise 0.5 when det F (FP64; J for J-fluid) exceeds jpc (jpc < 0: the
// material's default), else 0.
__global__ void __launch_bounds__(kThreads) k_particle_beta(State A, int64_t n, Lattice L, const DevMaterial *mats,
                                                            float dt, float jpc, BoundaryParticles bp,
                                                            Partition P) {
    for (int64_t k = blockIdx.x * (int64_t)kThreads + threadIdx.x; k < n; k += (int64_t)gridDim.x * kThreads) {
        const uint32_t i = (uint32_t)k;
        const uint32_t mw = chan_u(A.ch, i, kChMat);
        const DevMaterial mt = mats[mw & kMatIndexMask];
        const float h = dt >= 0.0f ? dt : scene_params(L, (uint32_t)mt.scene).dt;
        float vp[3], pos[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            vp[q] = chan(A.ch, i, kChV + q);
            pos[q] = chan(A.ch, i, kChX + q) + h * vp[q];
        }
        bool moving_in = false;
        for (int c = 0; c < L.ncols && !moving_in; ++c) {
            const DevCollider &col = L.cols[c];
            if (col.scene != mt.scene) continue;
            const float d[3] = {pos[0] - col.b[0], pos[1] - col.b[1], pos[2] - col.b[2]};
            const float *R = col.R;
            const float inv_s = 1.0f / col.s;
            float X[3], g[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) X[q] = (R[q] * d[0] + R[3 + q] * d[1] + R[6 + q] * d[2]) * inv_s;
            if (!(collider_phi(col, X, g) < 0.0f)) continue;
            float dot = 0.0f;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const float nq = R[3 * q] * g[0] + R[3 * q + 1] * g[1] + R[3 * q + 2] * g[2];
                dot += (vp[q] - col.dbdt[q]) * nq;
            }
            moving_in = dot < 0.0f;
        }
        if (!moving_in && bp.n > 0 && mt.scene == bp.scene) {
            const SceneParams sc = scene_params(L, (uint32_t)mt.scene);
            const double inv_scale = 1.0 / bnd_scale(bp);
            int base[3];
            bool inside = true;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const double c = floor((double)(pos[q] * sc.dxinv));
                inside = inside && c > -1.0e9 && c < 1.0e9;
                base[q] = inside ? (int)c : 0;
            }
            for (int o = 0; o < 8 && inside && !moving_in; ++o) {
                const int g[3] = {base[0] + (o >> 2), base[1] + ((o >> 1) & 1), base[2] + (o & 1)};
                const int lb[3] = {(g[0] >> 2) - L.origin, (g[1] >> 2) - L.origin, (g[2] >> 2) - L.origin};
                if (lb[0] < 0 || lb[1] < 0 || lb[2] < 0 || lb[0] >= sc.nb || lb[1] >= sc.nb || lb[2] >= sc.nb) continue;
                const int32_t gid = P.lat_gid[sc.lat_off + ((int64_t)lb[0] * sc.nb + lb[1]) * sc.nb + lb[2]];
                if (gid < 0) continue;
                const int cell = ((g[0] & 3) * 4 + (g[1] & 3)) * 4 + (g[2] & 3);
                const unsigned long long *a = bp.acc + (int64_t)gid * (kBndChannels * 64) + cell;
                const double bm = (double)(long long)a[0] * inv_scale;
                if (bm == 0.0) continue;
                double dot = 0.0, nn = 0.0;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const double vb = (double)(long long)a[(1 + q) * 64] * inv_scale / bm;
                    const double nq = (double)(long long)a[(4 + q) * 64] * inv_scale / bm;
                    dot += ((double)vp[q] - vb) * nq;
                    nn += nq * nq;
                }
                moving_in = nn > 0.0 && dot < 0.0;
            }
        }
        float beta = 0.0f;
        if (!moving_in) {
            double J;
            if (mw & kMatF) {
                double F[3][3];
#pragma unroll
                for (int q = 0; q < 9; ++q) F[q / 3][q % 3] = (double)chan(A.ch, i, kChF + q);
                J = F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) - F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0]) +
                    F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]);
            } else {
                J = (double)chan(A.ch, i, kChF);
            }
            // jpc < 0: the reference's JpcDefault, 1 for every material.  Its 1.1 branch
            // (Boundary.cu:389-391, 412-414, 466-468) tests the anisotropic-model variant,
            // variant<monostate, AnisotropicArap> (CUDA/Structures.hpp:30), for Cam-clay and never runs.
            const float jp = jpc >= 0.0f ? jpc : 1.0f;
            beta = J > (double)jp ? 0.5f : 0.0f;
        }
        chan(A.ch, i, kChBeta) = beta;
    }
}

// ------------------------------------------------------------------ particle volume
// ComputeParticleVolume (Maintenance.cu:17-56) with spatial_hashing's per-cell
// counts (projects/CUDA/Utils.hpp:770-843): the bucket of a particle is the cell
// floor(x / dx) (one bucket per grid cell).  That cell is the particle's base cell
// floor(x / dx - 1/2) or its +1 neighbour on each axis, so it lies in a grid block
// of the partition (blocks B + {0,1}^3), whose 64 scratch words count it.
__device__ __forceinline__ bool bucket_slot(const State &A, uint32_t i, const SceneParams &sc, const Lattice &L,
                                            const Partition &P, uint32_t &slot) {
    int lb[3], cl[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int c = (int)floorf(chan(A.ch, i, kChX + k) * sc.dxinv);
        lb[k] = (c >> 2) - L.origin;
        cl[k] = c & 3;
        if (lb[k] < 0 || lb[k] >= sc.nb) return false;
    }
    const int gb = P.lat_gid[(int64_t)sc.lat_off + ((int64_t)lb[0] * sc.nb + lb[1]) * sc.nb + lb[2]];
    if (gb < 0) return false;
    slot = (uint32_t)gb * 64u + (uint32_t)((cl[0] * 4 + cl[1]) * 4 + cl[2]);
    return true;
}

__global__ void __launch_bounds__(kThreads) k_cell_clear(Partition P, uint32_t *cnt) {
    const int64_t m = (int64_t)P.counts[1] * 64;
    for (int64_t k = blockIdx.x * (int64_t)kThreads + threadIdx.x; k < m; k += (int64_t)gridDim.x * kThreads)
        cnt[k] = 0u;
}

__global__ void __launch_bounds__(kThreads) k_cell_count(State A, int64_t n, Partition P, Lattice L,
                                                         const DevMaterial *mats, uint32_t *cnt, DevStatus *st) {
    bool bad = false;
    for (int64_t k = blockIdx.x * (int64_t)kThreads + threadIdx.x; k < n; k += (int64_t)gridDim.x * kThreads) {
        const uint32_t i = (uint32_t)k;
        const SceneParams sc = scene_params(L, (uint32_t)mats[chan_u(A.ch, i, kChMat) & kMatIndexMask].scene);
        uint32_t slot;
        if (bucket_slot(A, i, sc, L, P, slot)) atomicAdd(&cnt[slot], 1u);
        else bad = true;
    }
    if (bad) atomicOr(&st->bin_mismatch, 1u);
}

// vol = dx^3 / count and m = rho vol, written at the caller's particle index
// (cellVol = dx dx dx and density * vol, Maintenance.cu:41-52).
__global__ void __launch_bounds__(kThreads) k_cell_volume(State A, int64_t n, Partition P, Lattice L,
                                                          const DevMaterial *mats, const uint32_t *cnt, VolumeOut o) {
    for (int64_t k = blockIdx.x * (int64_t)kThreads + threadIdx.x; k < n; k += (int64_t)gridDim.x * kThreads) {
        const uint32_t i = (uint32_t)k;
        const DevMaterial mt = mats[chan_u(A.ch, i, kChMat) & kMatIndexMask];
        const SceneParams sc = scene_params(L, (uint32_t)mt.scene);
        uint32_t slot;
        if (!bucket_slot(A, i, sc, L, P, slot)) continue;  // flagged by k_cell_count
        const float v = ((sc.dx * sc.dx) * sc.dx) / (float)cnt[slot];
        const int64_t pid = chan_u(A.ch, i, kChPid);
        // the particle's own m and vol, which P2G and the stress read from now on (Maintenance.cu:50-51)
        chan(A.ch, i, kChVol) = v;
        chan(A.ch, i, kChMass) = mt.rho * v;
        if (o.vol) o.vol[view_index(o.layout, n, o.tile_channels, o.vol_chn, 1, 0, pid)] = v;
        if (o.mass) o.mass[view_index(o.layout, n, o.tile_channels, o.mass_chn, 1, 0, pid)] = mt.rho * v;
    }
}

// Per-scene max |v|^2 from the per-grid-block maxima (out zeroed by the caller):
// one atomic max on the float bits per wave whose blocks share a scene, per lane
// otherwise.  Non-negative floats order as their bits, and max is order
// independent, so the result is deterministic.
__global__ void __launch_bounds__(kThreads) k_reduce_vmax(const float *gvmax, const uint32_t *gblk_scene,
                                                          const uint32_t *counts, uint32_t *out) {
    const uint32_t ng = counts[1];
    for (int64_t g0 = blockIdx.x * (int64_t)kThreads; g0 < ng; g0 += (int64_t)gridDim.x * kThreads) {
        const int64_t g = g0 + threadIdx.x;
        const bool valid = g < ng;
        const uint32_t scene = valid ? gblk_scene[g] : 0xffffffffu;
        float v = valid ? gvmax[g] : 0.0f;
        const uint32_t s0 = __shfl(scene, 0, 64);
        if (__ballot(valid && scene != s0) == 0ull) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
            if ((threadIdx.x & 63) == 0 && v > 0.0f) atomicMax(&out[s0], __float_as_uint(v));
        } else if (valid && v > 0.0f) {
            atomicMax(&out[scene], __float_as_uint(v));
        }
    }
}

__global__ void __launch_bounds__(kThreads) k_set_steps(SceneParams *scenes, int32_t nscenes, StepArgs a, int mask) {
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= nscenes) return;
    SceneParams &sc = scenes[k];
    if (mask & SET_DT) sc.dt = a.dt;
    if (mask & SET_ACCEL) {
        sc.accel[0] = a.accel[0];
        sc.accel[1] = a.accel[1];
        sc.accel[2] = a.accel[2];
    }
    if (mask & SET_BOUNDARY) {
        sc.boundary = a.boundary;
        sc.pad = a.pad;
    }
}

inline unsigned grid_for(int64_t n, int64_t cap_blocks) {
    int64_t b = (n + kThreads - 1) / kThreads;
    if (b > cap_blocks) b = cap_blocks;
    if (b < 1) b = 1;
    return (unsigned)b;
}

// Launches over particle / grid blocks: `max_blocks` is the host's estimate of
// the block count (the kernels loop over any blocks beyond the grid).
inline unsigned persistent_grid(int64_t max_blocks, int64_t per_wg, int64_t cap) {
    int64_t g = (max_blocks + per_wg - 1) / per_wg;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (unsigned)g;
}

}  // namespace

__global__ void __launch_bounds__(kThreads) k_extract_pid(const float *ch, int64_t n, uint32_t *out) {
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
        out[i] = chan_u(ch, (uint32_t)i, kChPid);
}

void launch_extract_pid(State A, int64_t n, uint32_t *out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_extract_pid, dim3(grid_for(n, 8192)), dim3(kThreads), 0, s, (const float *)A.ch, n, out);
}

void launch_bin_keys(const float *x, int64_t stride, int64_t n, const DevMaterial *mats, Lattice L, uint32_t *keys,
                     DevStatus *st, hipStream_t s, bool check_sorted) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_bin_keys, dim3(grid_for(n, 8192)), dim3(kThreads), 0, s, x, stride, n, mats, L, keys, st,
                       check_sorted ? 1 : 0);
}

void launch_set_steps(SceneParams *scenes, int32_t nscenes, StepArgs a, int mask, hipStream_t s) {
    if (nscenes <= 0 || mask == 0) return;
    hipLaunchKernelGGL(k_set_steps, dim3((unsigned)((nscenes + kThreads - 1) / kThreads)), dim3(kThreads), 0, s,
                       scenes, nscenes, a, mask);
}

void launch_partition_from_keys(const uint32_t *skeys, int64_t n, Lattice L, Partition P, Binning B, State sorted,
                                const DevMaterial *mats, DevStatus *st, hipStream_t s) {
    // lat_flag and dcount are zero here: every scan clears the entries it was given to read
    hipLaunchKernelGGL(k_scan_reset, dim3(grid_for(B.scan_tiles, 1024)), dim3(kThreads), 0, s, B, 1u);
    if (n > 0)
        hipLaunchKernelGGL(k_mark_blocks, dim3(grid_for(n, 8192)), dim3(kThreads), 0, s, skeys, n, L,
                           (const float *)sorted.ch, mats, B.lat_flag, B.dcount);
    // the scan forms the block starts from the counts k_mark_blocks left in dcount
    hipLaunchKernelGGL(k_lattice_scan, dim3((unsigned)B.scan_tiles), dim3(kScanThreads), 0, s, L, P, B,
                       (const uint32_t *)P.tile_used, st);
}

void launch_gather_state(State src, int64_t src_cap, State dst, int64_t dst_cap, const uint32_t *perm, int64_t n,
                         hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather_state, dim3(grid_for(n, 8192)), dim3(kThreads), 0, s, src, src_cap, dst, dst_cap,
                       perm, n);
}

void launch_stress(State A, int64_t n, Lattice L, const DevMaterial *mats, float *sig, int64_t stride,
                   hipStream_t s, bool general) {
    if (n <= 0) return;
    if (general)
        hipLaunchKernelGGL(k_stress<true>, dim3(grid_for(n, 2048)), dim3(kThreads), 0, s, A, n, L, mats, sig, stride);
    else
        hipLaunchKernelGGL(k_stress<false>, dim3(grid_for(n, 2048)), dim3(kThreads), 0, s, A, n, L, mats, sig, stride);
}

// k_p2g<MODE> with or without the particles' own mass and volume (L.pp_mass)
template <int MODE>
void launch_p2g_mode(dim3 g, hipStream_t s, State A, Partition P, Lattice L, const DevMaterial *mats, const float *sig,
                     int64_t stride, float *tiles, DevStatus *st, AdjScatter adj, uint32_t skip_upto) {
    if (L.pp_mass)
        hipLaunchKernelGGL((k_p2g<MODE, true>), g, dim3(kP2GThreads), 0, s, A, P, L, mats, sig, stride, tiles, st, adj,
                           skip_upto);
    else
        hipLaunchKernelGGL((k_p2g<MODE, false>), g, dim3(kP2GThreads), 0, s, A, P, L, mats, sig, stride, tiles, st,
                           adj, skip_upto);
}

void launch_p2g(State A, Partition P, Lattice L, const DevMaterial *mats, const float *sig, int64_t stride,
                float *tiles, DevStatus *st, int64_t max_blocks, hipStream_t s, bool sparse, bool flush) {
    if (max_blocks <= 0) return;
    const dim3 g(persistent_grid(max_blocks, 1, 1 << 20));
    // with sparse, blocks of <= kSparseMax particles go to the one-wave kernel (one chunk either way)
    const uint32_t skip = sparse ? (uint32_t)kSparseMax : 0u;
    if (!flush)
        launch_p2g_mode<kP2GForward>(g, s, A, P, L, mats, sig, stride, tiles, st, AdjScatter{}, skip);
    else if (L.pp_mass)
        hipLaunchKernelGGL((k_p2g<kP2GForward, true, true>), g, dim3(kP2GThreads), 0, s, A, P, L, mats, sig, stride,
                           tiles, st, AdjScatter{}, skip);
    else
        hipLaunchKernelGGL((k_p2g<kP2GForward, false, true>), g, dim3(kP2GThreads), 0, s, A, P, L, mats, sig, stride,
                           tiles, st, AdjScatter{}, skip);
    if (sparse)
        hipLaunchKernelGGL(k_p2g_sparse, dim3(persistent_grid(max_blocks, 4, 1 << 20)), dim3(256), 0, s, A, P, L,
                           mats, sig, stride, tiles, st);
}

void launch_p2g_aflip(State A, Partition P, Lattice L, const DevMaterial *mats, const float *sig, int64_t stride,
                      float *tiles, float *tiles2, DevStatus *st, int64_t max_blocks, hipStream_t s) {
    if (max_blocks <= 0) return;
    const dim3 g(persistent_grid(max_blocks, 1, 1 << 20));
    launch_p2g_mode<kP2GMomentum>(g, s, A, P, L, mats, sig, stride, tiles, st, AdjScatter{}, 0u);
    launch_p2g_mode<kP2GStress>(g, s, A, P, L, mats, sig, stride, tiles2, st, AdjScatter{}, 0u);
}

void launch_p2g_flip(State A, Partition P, Lattice L, const DevMaterial *mats, const float *sig, int64_t stride,
                     float *tiles, float *tiles2, DevStatus *st, int64_t max_blocks, hipStream_t s) {
    if (max_blocks <= 0) return;
    const dim3 g(persistent_grid(max_blocks, 1, 1 << 20));
    launch_p2g_mode<kP2GFlipMomentum>(g, s, A, P, L, mats, sig, stride, tiles, st, AdjScatter{}, 0u);
    launch_p2g_mode<kP2GFlipStress>(g, s, A, P, L, mats, sig, stride, tiles2, st, AdjScatter{}, 0u);
}

void launch_grid_aflip(int mode, Partition P, Lattice L, const float *tiles, const float *tiles2, float *grid,
                       float *gstar, float *gvmax, int64_t max_blocks, hipStream_t s, uint32_t *vclear,
                       int32_t nclear) {
    if (max_blocks <= 0) return;
    hipLaunchKernelGGL(k_grid_aflip, dim3(persistent_grid(max_blocks, 4, 1 << 20)), dim3(kThreads), 0, s, mode, P, L,
                       tiles, tiles2, grid, gstar, gvmax, vclear, nclear);
}

void launch_p2g_adjoint(State A, Partition P, Lattice L, AdjScatter adj, float *tiles, DevStatus *st,
                        int64_t max_blocks, hipStream_t s) {
    if (max_blocks <= 0) return;
    hipLaunchKernelGGL(k_p2g<kP2GAdjoint>, dim3(persistent_grid(max_blocks, 1, 1 << 20)), dim3(kP2GThreads), 0, s, A,
                       P, L, (const DevMaterial *)nullptr, (const float *)nullptr, (int64_t)0, tiles, st, adj, 0u);
}

void launch_grid(int mode, Partition P, Lattice L, const float *tiles, float *grid, float *gvmax,
                 int64_t max_blocks, hipStream_t s, float *out, uint32_t *vclear, int32_t nclear) {
    if (max_blocks <= 0) return;
    hipLaunchKernelGGL(k_grid, dim3(persistent_grid(max_blocks, 4, 1 << 20)), dim3(kThreads), 0, s, mode, P, L, tiles,
                       grid, gvmax, out ? out : grid, vclear, nclear);
}

void launch_reduce_vmax(const float *gvmax, Partition P, int64_t max_blocks, uint32_t *out, hipStream_t s) {
    if (max_blocks <= 0) return;
    hipLaunchKernelGGL(k_reduce_vmax, dim3(grid_for(max_blocks, 4096)), dim3(kThreads), 0, s, gvmax,
                       (const uint32_t *)P.gblk_scene, (const uint32_t *)P.counts, out);
}

void launch_g2p_move(State A, Partition P, Lattice L, const float *grid, const float *gvmax, Binning B,
                     DevStatus *st, int64_t max_blocks, hipStream_t s, int wg, const float *gstar, bool nonaffine,
                     bool fused) {
    if (max_blocks <= 0) return;
    const dim3 g(persistent_grid(max_blocks, 1, 1 << 20));
    const int na = nonaffine ? 1 : 0;
    if (gstar)
        hipLaunchKernelGGL((k_g2p_move<64, true>), g, dim3(64), 0, s, A, P, L, grid, gvmax, B, st, gstar, na);
    else if (fused && wg == 64)
        hipLaunchKernelGGL((k_g2p_move<64, false, true>), g, dim3(64), 0, s, A, P, L, grid, gvmax, B, st, gstar, na);
    else if (fused)
        hipLaunchKernelGGL((k_g2p_move<kThreads, false, true>), g, dim3(kThreads), 0, s, A, P, L, grid, gvmax, B, st,
                           gstar, na);
    else if (wg == 64)
        hipLaunchKernelGGL((k_g2p_move<64, false>), g, dim3(64), 0, s, A, P, L, grid, gvmax, B, st, gstar, na);
    else
        hipLaunchKernelGGL((k_g2p_move<kThreads, false>), g, dim3(kThreads), 0, s, A, P, L, grid, gvmax, B, st,
                           gstar, na);
}

void launch_g2p2g(State A, State Bst, Partition P, Partition Pn, Lattice L, const float *grid, Binning Bn,
                  uint32_t *arr, float *tiles, const DevMaterial *mats, DevStatus *st, int aux, bool write_vc,
                  bool general, int64_t n, int64_t max_blocks, hipStream_t s) {
    if (max_blocks <= 0 || n <= 0) return;
    FusedArgs u{};
    u.src = A.ch;
    u.dst = Bst.ch;
    u.pblk_start = P.pblk_start;
    u.lat_pid = P.lat_pid;
    u.lat_gid = P.lat_gid;
    u.n_counts = Pn.counts;
    u.n_pblk_lat = Pn.pblk_lat;
    u.n_pblk_start = Pn.pblk_start;
    u.n_pblk_scene = Pn.pblk_scene;
    u.cnt27 = Bn.cnt27;
    u.stay = Bn.code;
    u.mv = Bn.mvlist;
    u.arr = arr;
    u.grid = grid;
    u.tiles = tiles;
    u.mats = mats;
    u.L = L;
    u.st = st;
    u.aux = aux;
    u.write_vc = write_vc ? 1 : 0;
    u.n = (uint32_t)n;
    const dim3 g(persistent_grid(max_blocks, 1, 1 << 20));
    if (general) {
        if (L.pp_mass) hipLaunchKernelGGL((k_g2p2g<true, true>), g, dim3(kP2GThreads), 0, s, u);
        else hipLaunchKernelGGL((k_g2p2g<false, true>), g, dim3(kP2GThreads), 0, s, u);
    } else {
        if (L.pp_mass) hipLaunchKernelGGL((k_g2p2g<true, false>), g, dim3(kP2GThreads), 0, s, u);
        else hipLaunchKernelGGL((k_g2p2g<false, false>), g, dim3(kP2GThreads), 0, s, u);
    }
}

void launch_morton_codes(int32_t layout, int32_t nch, int64_t n, const float *src, int32_t x_chn, const float lo[3],
                         const float ext[3], uint32_t *codes, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_morton_codes, dim3(grid_for(n, 8192)), dim3(kThreads), 0, s, layout, nch, n, src, x_chn, lo[0],
                       lo[1], lo[2], ext[0], ext[1], ext[2], codes);
}

void launch_permute_channels(int32_t layout, int32_t nch, int64_t n, const float *src, float *dst,
                             const uint32_t *order, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_permute_channels, dim3(grid_for(n, 8192)), dim3(kThreads), 0, s, layout, nch, n, src, dst, order);
}

void launch_lattice_scan(Lattice L, Partition Pn, Binning B, const uint32_t *prev_used, int64_t n, DevStatus *st,
                         hipStream_t s) {
    (void)n;
    hipLaunchKernelGGL(k_lattice_scan, dim3((unsigned)B.scan_tiles), dim3(kScanThreads), 0, s, L, Pn, B, prev_used,
                       st);
}

void launch_g2p_update(bool scatter, State A, State Bst, Partition P, Partition Pn, Lattice L, const float *grid,
                       Binning Bn, int64_t max_blocks, hipStream_t s, int wg, const float *gstar, int aux,
                       bool nonaffine) {
    if (max_blocks <= 0) return;
    UpdateArgs u{};
    u.src = A.ch;
    u.dst = scatter ? Bst.ch : A.ch;
    u.counts = P.counts;
    u.pblk_lat = P.pblk_lat;
    u.pblk_start = P.pblk_start;
    u.pblk_scene = P.pblk_scene;
    u.parts = part_view(P, L.part_shift);
    u.lat_gid = P.lat_gid;
    u.lat_pid = P.lat_pid;
    u.lat_xf = P.lat_xf;
    u.n_lat_pid = Pn.lat_pid;
    u.n_pblk_start = Pn.pblk_start;
    u.cnt27 = Bn.cnt27;
    u.code = Bn.code;
    u.grid = grid;
    u.L = L;
    u.gstar = gstar;
    u.aux = aux;
    u.nonaffine = nonaffine ? 1 : 0;
    const unsigned g = persistent_grid(max_blocks, 1, 1 << 20);
    if (gstar && scatter)
        hipLaunchKernelGGL((k_g2p_update<true, 64, true>), dim3(g), dim3(64), 0, s, u);
    else if (gstar)
        hipLaunchKernelGGL((k_g2p_update<false, 64, true>), dim3(g), dim3(64), 0, s, u);
    else if (scatter && wg == 64)
        hipLaunchKernelGGL((k_g2p_update<true, 64>), dim3(g), dim3(64), 0, s, u);
    else if (scatter)
        hipLaunchKernelGGL((k_g2p_update<true, kThreads>), dim3(g), dim3(kThreads), 0, s, u);
    else
        hipLaunchKernelGGL((k_g2p_update<false, kThreads>), dim3(g), dim3(kThreads), 0, s, u);
}

void launch_import(State dst, int64_t cap, const void *view, int64_t off, const DevMaterial *mats, int32_t nmat,
                   DevStatus *st, hipStream_t s) {
    const zmpm_particles v = *(const zmpm_particles *)view;
    if (v.n <= 0) return;
    hipLaunchKernelGGL(k_import, dim3(grid_for(v.n, 8192)), dim3(kThreads), 0, s, dst, cap, v, off, mats, nmat, st);
}

void launch_colliders(Partition P, Lattice L, float *grid, float *gvmax, int64_t max_blocks, bool vmax,
                      hipStream_t s) {
    if (L.ncols <= 0 || max_blocks <= 0) return;
    hipLaunchKernelGGL(k_colliders, dim3(persistent_grid(max_blocks, 4, 1 << 20)), dim3(kThreads), 0, s, P, L, grid,
                       gvmax, vmax ? 1 : 0);
}

void launch_step_colliders(Lattice L, float dt, hipStream_t s) {
    if (L.ncols <= 0) return;
    hipLaunchKernelGGL(k_step_colliders, dim3(1), dim3(kMaxColliders), 0, s, L, dt);
}

void launch_return_mapping(State A, int64_t n, const DevMaterial *mats, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_return_mapping, dim3(grid_for(n, 8192)), dim3(kThreads), 0, s, A, n, mats);
}

void launch_boundary_particles(BoundaryParticles bp, Partition P, Lattice L, float *grid, float *gvmax,
                               int64_t max_gblocks, hipStream_t s, bool apply) {
    if (bp.n <= 0 || max_gblocks <= 0) return;
    hipLaunchKernelGGL(k_bnd_clear, dim3(grid_for(max_gblocks * kBndChannels * 64, 8192)), dim3(kThreads), 0, s, bp,
                       P);
    hipLaunchKernelGGL(k_bnd_scale, dim3(grid_for(bp.n, 1024)), dim3(kThreads), 0, s, bp);
    hipLaunchKernelGGL(k_bnd_p2g, dim3(grid_for(bp.n, 8192)), dim3(kThreads), 0, s, bp, P, L);
    if (apply)
        hipLaunchKernelGGL(k_bnd_apply, dim3(persistent_grid(max_gblocks, 4, 1 << 20)), dim3(kThreads), 0, s, bp, P,
                           L, grid, gvmax);
}

void launch_init_aux(State A, int64_t n, const DevMaterial *mats, int mask, hipStream_t s) {
    if (n <= 0 || mask == 0) return;
    hipLaunchKernelGGL(k_init_aux, dim3(grid_for(n, 8192)), dim3(kThreads), 0, s, A, n, mats, mask);
}

void launch_particle_beta(State A, int64_t n, Lattice L, const DevMaterial *mats, float dt, float jp_critical,
                          BoundaryParticles bp, Partition P, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_particle_beta, dim3(grid_for(n, 8192)), dim3(kThreads), 0, s, A, n, L, mats, dt,
                       jp_critical, bp, P);
}

void launch_particle_volume(State A, int64_t n, Partition P, Lattice L, const DevMaterial *mats, uint32_t *cnt,
                            int64_t max_gblocks, VolumeOut out, DevStatus *st, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cell_clear, dim3(grid_for(std::max<int64_t>(max_gblocks, 1) * 64, 8192)), dim3(kThreads), 0,
                       s, P, cnt);
    hipLaunchKernelGGL(k_cell_count, dim3(grid_for(n, 8192)), dim3(kThreads), 0, s, A, n, P, L, mats, cnt, st);
    hipLaunchKernelGGL(k_cell_volume, dim3(grid_for(n, 8192)), dim3(kThreads), 0, s, A, n, P, L, mats,
                       (const uint32_t *)cnt, out);
}

void launch_export(State src, int64_t n_res, const void *view, int64_t off, const DevMaterial *mats, int aux,
                   hipStream_t s) {
    const zmpm_particles v = *(const zmpm_particles *)view;
    if (v.n <= 0 || n_res <= 0) return;
    hipLaunchKernelGGL(k_export, dim3(grid_for(n_res, 8192)), dim3(kThreads), 0, s, src, n_res, v, off, mats, aux);
}

}  // namespace zmpm
