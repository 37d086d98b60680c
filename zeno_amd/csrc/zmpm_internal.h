// Internal declarations shared by the libzmpm translation units (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "lookback.h"

namespace zmpm {

constexpr int kBlockSide = 4;        // grid block = 4^3 cells (ZenoGrid::grid_t side_length, Structures.hpp:595)
constexpr int kBlockCells = 64;
constexpr int kTileSide = 6;         // P2G tile of one particle block: nodes [K, K+6)^3
constexpr int kTileNodes = 216;
// Parts: the lattice scan cuts a particle block of d > kPartParticles particles into
// ceil(d / kPartParticles) parts of consecutive particles (the last takes the rest), each its
// own workgroup in P2G (its own tile, the grid kernel summing a block's parts in order) and in
// both G2P passes, so that a block of thousands of particles does not set a launch's length
// alone.  A block's first part is the block itself (its id indexes tiles and counts as
// before); the extra parts are numbered from the top of the part-indexed arrays down, and
// kernels look them up only when the partition has any.  kNoSplit keeps one part per block
// (the fused G2P / P2G pass works per block).
constexpr uint32_t kPartShift = 10;
constexpr uint32_t kPartParticles = 1u << kPartShift;  // 1,024
constexpr uint32_t kNoSplit = 31;                      // a part shift no block reaches
// lattice scan: the particle-block count's bits in the packed (particles | blocks) prefix
constexpr int kPBlkBits = 27;
constexpr int kGridChannels = 4;     // m, v0, v1, v2 (MakeZSGrid apic channels, Generation.cu:1490)
// Resident particle channels: x3 v3 C9 F9 (J-fluid: F0 = J), material index and
// caller index (both as uint32 bits), the ASFLIP beta, the Cam-clay logJp and
// the per-particle Lame parameters mu, lam, the ARAP fiber, and the per-particle
// mass and volume.
constexpr int kChX = 0, kChV = 3, kChC = 6, kChF = 15, kChMat = 24, kChPid = 25, kChBeta = 26, kChLogJp = 27;
constexpr int kChMu = 28, kChLam = 29, kChFiber = 30, kChMass = 33, kChVol = 34;
constexpr int kTileCh = 35;
// Channels the G2P scatter carries only while something reads them.
constexpr int kAuxBeta = 1;   // AFLIP / FLIP transfers
constexpr int kAuxLogJp = 2;  // Cam-clay materials
constexpr int kAuxLame = 4;   // per-particle mu, lam (imported with the particles)
constexpr int kAuxFiber = 8;  // the ARAP fiber direction a (3 channels)
constexpr int kAuxMass = 16;  // per-particle mass and volume (imported, or written by the particle volume)
constexpr float kLogJpInit = -0.04f;  // ToZSParticles' initial logJp (Generation.cpp:987)
// Material word: table index in the low 16 bits, bit 31 set for fixed-corotated particles.
constexpr uint32_t kMatIndexMask = 0xffffu;
constexpr uint32_t kMatF = 1u << 31;        // F-based stress (fixed corotated or StVK-Hencky) from k_stress
constexpr uint32_t kMatPlastic = 1u << 30;  // return mapping after G2P
constexpr int kTileParticles = 64;
constexpr int kDirs = 27;            // a particle moves to one of the 3^3 neighbour blocks per substep

// Material record as the kernels read it (derived on the host from zmpm_material).
struct DevMaterial {
    int32_t model;
    float mu, lam;   // Lame parameters (FCR)
    float mass;      // rho * vol
    float vol;
    float bulk;      // K (J-fluid)
    int32_t scene;   // scene of the particles using this material
    float rho;       // density
    int32_t plasticity;  // ZMPM_PLASTIC_*
    float dp_alpha;      // Drucker-Prager: sqrt(2/3) 2 sin(phi) / (3 - sin(phi))
    float yield_stress;  // von Mises
    float flip_beta;     // ASFLIP positional adjustment: initial value of the particles' beta
    float nacc_M;        // Cam-clay critical-state slope from the friction angle
    float nacc_beta;     // Cam-clay cohesion ratio
    float nacc_xi;       // Cam-clay hardening
    float nacc_bulk;     // Cam-clay: bulk modulus 2 mu / 3 + lambda of the StVK-Hencky strain
    float aniso_mu;      // ARAP fiber stiffness strength * mu (0: no fiber term)
};

// Per-scene parameters (device table).  Scene s owns lattice entries
// [lat_off, lat_off + nb^3) of the concatenated block lattice, origin block -1
// on every axis.
struct SceneParams {
    float dx, dxinv, dt, Dinv;  // Dinv = 4 / dx^2 (quadratic B-spline)
    float accel[3];
    int32_t boundary, pad, n_cells;
    int32_t lat_off, nb;
    float pad0, pad1;
};

// Block lattice: the scenes' dense maps over block coordinates [origin, origin + nb)
// per axis, one after another.
// A level-set collider on the device: the caller's zmpm_collider with the
// rotation as a matrix (local -> world).
struct DevCollider {
    int32_t type, shape, scene, pad0;
    int32_t dims[3];
    float spacing;
    float shape_params[6];
    float origin[3];
    float pad1;
    const float *sdf;
    float b[3], dbdt[3];
    float s, dsdt;
    float R[9];
    float omega[3];
    float pad2;
    // SDF_LEAVES: leaf id (or -1) of every 8^3 box of the leaves' bounding box, boxes
    // [leaf_lo, leaf_lo + dims) per axis, row-major (x slowest)
    const int32_t *leaf_map;
    int32_t leaf_lo[3];
    int32_t pad3;
};
constexpr int kMaxColliders = 64;
constexpr int32_t kMaxSdfLeaves = 1 << 24;     // leaves of one SDF_LEAVES collider
constexpr int64_t kMaxSdfLeafMap = 1ll << 28;  // 8^3 boxes of its leaves' bounding box

struct Lattice {
    int32_t origin;    // first block coordinate (in blocks) on every axis
    int32_t nscenes;
    int64_t size;      // sum of nb^3 over the scenes
    const SceneParams *scenes;  // device table
    SceneParams s0;    // the table's entry 0 by value: one-scene contexts read no table
    DevCollider *cols; // device table of level-set colliders
    int32_t ncols;
    int32_t pp_lame;   // the stress reads each particle's own mu, lam channels (Transfer.cu:861-864)
    int32_t pp_mass;   // P2G and the stress read each particle's own m, vol channels (Transfer.cu:857-858)
    int32_t key_shift; // binning keys: lattice index << 6 | cell (6), or the lattice index alone (0)
    uint32_t part_shift; // parts of 2^part_shift particles (kNoSplit: one part per block)
};

// Resident particle state: tiles of 64 particles, kTileCh channels per tile,
// channel c of particle i at ch[((i / 64) * kTileCh + c) * 64 + i % 64].
struct State {
    float *ch;
};
inline int64_t state_floats(int64_t cap) { return (cap + kTileParticles - 1) / kTileParticles * kTileParticles * kTileCh; }

// Partition: particle blocks (sorted by lattice index, particles stably grouped
// by block) and grid blocks B + {0,1}^3.
struct Partition {
    int32_t *lat_pid;     // lattice -> particle block id or -1
    int32_t *lat_gid;     // lattice -> grid block id or -1
    uint32_t *pblk_lat;   // particle block -> lattice index
    uint32_t *pblk_start; // particle block -> first particle (np + 1 entries)
    uint32_t *gblk_lat;   // grid block -> lattice index
    uint32_t *pblk_scene; // particle block -> scene
    uint32_t *gblk_scene; // grid block -> scene
    uint32_t *tile_used;  // per lattice-scan tile: 1 if the last scan into this buffer marked an entry there
    uint32_t *counts;     // [0] = particle blocks, [1] = grid blocks, [2] = extra parts
    uint32_t *pblk_x;     // particle block -> extra parts before it (np + 1 entries)
    uint8_t *lat_xf;      // lattice -> 1 if the particle block there has extra parts
    uint32_t *part_blk;   // extra part k at [parts_cap - 1 - k] -> its particle block
    uint32_t *part_lo;    // extra part k at [parts_cap - 1 - k] -> its first particle
    uint32_t parts_cap;   // entries of the part-indexed arrays (tiles, move counts)
};

// Scratch shared by the binning kernels of one context.  lat_flag and dcount
// are all zero between substeps (the lattice scan clears what it reads).
struct Binning {
    uint32_t *lat_flag;   // per lattice entry: bit0 particle block, bit1 grid block
    uint32_t *dcount;     // per lattice entry: particles moving into that block
    uint32_t *cnt27;      // per part of the old partition: particles leaving in each of the 27 directions
    uint32_t *code;       // per particle (old position): (stable rank in its (block, direction) group << 5) | direction;
                          // for the fused pass, the block's stayers in order at [block's first particle + rank]
    uint32_t *mvlist;     // fused pass: the block's movers in order at [block's first particle + k], (index << 5) | dir
    uint32_t *tile_mark;  // per lattice-scan tile: 1 forces the next scan to visit it (full re-bin)
    uint32_t *vmax_out;   // per-scene MaxVelSqr words the move pass reduces into (k_grid clears them)
    int32_t mark_radius;  // scan tiles between a block's tile and the farthest entry it can mark
    ScanState scan_p;     // look-back chain over (particle blocks | particle counts)
    ScanState scan_g;     // look-back chain over grid blocks
    int64_t scan_tiles;   // tiles of the lattice scan
    uint32_t *host_counts;  // host-mapped copy of the latest partition's block counts (launch sizing)
};

// Device-side status words.
struct DevStatus {
    uint32_t unused0;
    uint32_t domain_error;      // particles outside the lattice or not finite
    uint32_t bad_material;      // material index outside the table
    uint32_t cfl_error;         // a particle moved more than one block in one substep
    uint32_t bin_mismatch;      // a particle's cell lies outside its block (must never happen)
    uint32_t scan_stall;        // a look-back wait exceeded its bound (must never happen)
    uint32_t pad[2];
};

// ---- scan / sort (sort.hip)
constexpr int64_t kSortChunk = 2048;  // keys per radix-sort tile (256 threads x 8)
constexpr int64_t kSortRadix = 256;   // 8-bit digits
constexpr size_t align_up256(size_t b) { return (b + 255) & ~size_t(255); }
// bytes of exclusive_scan_u32's scratch for n values: one partial per tile, plus one
constexpr size_t scan_scratch_bound(int64_t n) {
    return align_up256(sizeof(uint32_t) * (size_t)((n + kSortChunk - 1) / kSortChunk + 1));
}
// bytes of radix_sort_pairs' scratch for n keys: one ping-pong pair of key and value
// arrays (the outputs are the other), the per-tile digit histograms and their offsets,
// and the scan of the histograms
constexpr size_t radix_scratch_bound(int64_t n) {
    return 2 * align_up256(sizeof(uint32_t) * (size_t)n) +
           2 * align_up256(sizeof(uint32_t) * (size_t)(((n + kSortChunk - 1) / kSortChunk) * kSortRadix)) +
           scan_scratch_bound(((n + kSortChunk - 1) / kSortChunk) * kSortRadix);
}
size_t scan_scratch_bytes(int64_t n);
// out[i] = sum_{j<i} ((in[j] >> shift) & mask)
void exclusive_scan_u32(const uint32_t *in, uint32_t *out, int64_t n, void *scratch, hipStream_t s, int shift,
                        uint32_t mask);
size_t radix_scratch_bytes(int64_t n);
// Stable LSD radix sort by bits [bit_lo, bit_hi) of keys.  vals_in == nullptr means vals = iota.
// Result in keys_out/vals_out (also used between passes, so they must not alias the inputs);
// inv_out (optional) receives inv[vals[k]] = k.  Launches kernels only: no allocation, no memset.
void radix_sort_pairs(const uint32_t *keys_in, const uint32_t *vals_in, int64_t n, int bit_lo, int bit_hi,
                      uint32_t *keys_out, uint32_t *vals_out, uint32_t *inv_out, void *scratch, hipStream_t s);
void fill_u32(uint32_t *p, uint32_t value, int64_t n, hipStream_t s);

// ---- MPM kernels (mpm.hip)
// Step parameters applied to every scene by launch_set_steps (mask bits below).
struct StepArgs {
    float dt;
    float accel[3];
    int32_t boundary, pad;
};
enum { SET_DT = 1, SET_ACCEL = 2, SET_BOUNDARY = 4 };
void launch_set_steps(SceneParams *scenes, int32_t nscenes, StepArgs a, int mask, hipStream_t s);

// x: SoA planes x[k * stride + i], or the resident tiled state when stride < 0.
// Scene of each particle: mats[material].scene for the resident state; scene 0 for SoA input.
// check_sorted: status->bin_mismatch if the particles' blocks are not in nondecreasing order
void launch_bin_keys(const float *x, int64_t stride, int64_t n, const DevMaterial *mats, Lattice L, uint32_t *keys,
                     DevStatus *st, hipStream_t s, bool check_sorted = false);
// out[i] = caller index of resident particle i
void launch_extract_pid(State A, int64_t n, uint32_t *out, hipStream_t s);
// Partition of a key array sorted by block (the full re-bin path); `sorted` is the
// state already permuted into that order (for the per-block material flags).
void launch_partition_from_keys(const uint32_t *skeys, int64_t n, Lattice L, Partition P, Binning B, State sorted,
                                const DevMaterial *mats, DevStatus *st, hipStream_t s);
// number of scenes a launch may address through SceneParams::lat_off lookups
constexpr int kMaxScenes = 1 << 15;
constexpr int kMaxMaterials = 1 << 12;  // fits the 16-bit index of the material word
void launch_gather_state(State src, int64_t src_cap, State dst, int64_t dst_cap, const uint32_t *perm, int64_t n,
                         hipStream_t s);
enum { kP2GForward = 0, kP2GAdjoint = 1, kP2GMomentum = 2, kP2GStress = 3, kP2GFlipMomentum = 4, kP2GFlipStress = 5 };
// Per-particle payload of the adjoint transfer (planes of `stride` entries, resident order):
// a[3] and B[3][3] such that node i receives W (a + B (x_i - x_p)).
struct AdjScatter {
    const float *a;
    const float *B;
    int64_t stride;
};
// Adjoint of G2P's velocity gather: tiles receive sum_p W (a_p + B_p d) per node (mass channel 0).
void launch_p2g_adjoint(State A, Partition P, Lattice L, AdjScatter adj, float *tiles, DevStatus *st,
                        int64_t max_blocks, hipStream_t s);
// Adjoint arrays of a differentiated substep, by caller particle index: planes of n
// entries x (3), v (3), C (9), F (9; J-fluid: F[0] = dL/dJ); mu, lam accumulate dL/dmu, dL/dlam.
struct AdjArrays {
    float *x, *v, *C, *F, *mu, *lam;
    int64_t n;
};
void launch_g2p_adjoint(State A, Partition P, Lattice L, const float *gridv, AdjArrays adj, float *pay_a, float *pay_B,
                        int64_t stride, int64_t max_blocks, hipStream_t s);
void launch_grid_adjoint(Partition P, Lattice L, const float *grid, float *gbar, int64_t max_blocks, hipStream_t s);
void launch_p2g_gather_adjoint(State A, Partition P, Lattice L, const DevMaterial *mats, const float *gbar,
                               AdjArrays adj, int64_t max_blocks, hipStream_t s);
// Stress term sig = -dt D^-1 vol P(F) F^T of every F-based particle (planes of stride entries).
// general: some F-based material is not plain fixed corotated (another model or a fiber term).
void launch_stress(State A, int64_t n, Lattice L, const DevMaterial *mats, float *sig, int64_t stride,
                   hipStream_t s, bool general = true);
// P2G of every particle block into its tile (fixed-corotated stress terms from launch_stress).
// sparse: blocks of at most 64 particles go to a one-wave-per-block kernel (a second launch).
// flush: the fused pass's chunk-flush grouping of blocks of more than one chunk (k_p2g FLUSH)
void launch_p2g(State A, Partition P, Lattice L, const DevMaterial *mats, const float *sig, int64_t stride,
                float *tiles, DevStatus *st, int64_t max_blocks, hipStream_t s, bool sparse = false,
                bool flush = false);
// mode bits for launch_grid
enum { GRID_COMBINE = 1, GRID_UPDATE = 2, GRID_BOUNDARY = 4, GRID_CLEAR = 8, GRID_KEEP = 16 };
// gvmax: per grid block max |v|^2, written by GRID_UPDATE
// vclear (optional): nclear words the first workgroup sets to zero (the substep's MaxVelSqr
// output, for the launch_reduce_vmax that follows: no memset in the substep)
void launch_grid(int mode, Partition P, Lattice L, const float *tiles, float *grid, float *gvmax,
                 int64_t max_blocks, hipStream_t s, float *out = nullptr, uint32_t *vclear = nullptr,
                 int32_t nclear = 0);
// out[scene] = max over the scene's grid blocks of gvmax, as float bits (out zeroed first)
void launch_reduce_vmax(const float *gvmax, Partition P, int64_t max_blocks, uint32_t *out, hipStream_t s);
// G2P pass 1 of the fused substep: direction of each particle's new block, per-block
// direction counts, destination marks.  Also resets the lattice scan's state.
// wg: threads per workgroup, 64 or 256 (one particle block per workgroup either way)
// gstar (with mats): the affine FLIP transfers' vstar grid (nullptr: APIC).
// nonaffine: the FLIP transfer (x += dt vstar, no beta term).
// fused: the stayer and mover lists of the fused pass in place of the codes (APIC only)
void launch_g2p_move(State A, Partition P, Lattice L, const float *grid, const float *gvmax, Binning B, DevStatus *st,
                     int64_t max_blocks, hipStream_t s, int wg = 256, const float *gstar = nullptr,
                     bool nonaffine = false, bool fused = false);
// The fused pass (APIC): G2P of this substep from A (partition P, grid) into B in the new
// partition Pn's order, and the next substep's stress and P2G into tiles (per block of Pn),
// bit-identical to launch_g2p_update + launch_stress + launch_p2g.  arr: n entries of scratch.
// write_vc: also store v and C (otherwise left unwritten: only P2G reads them, from registers here).
// general: a solid model other than fixed corotated, plasticity or the ARAP fiber term is in use.
void launch_g2p2g(State A, State B, Partition P, Partition Pn, Lattice L, const float *grid, Binning Bn,
                  uint32_t *arr, float *tiles, const DevMaterial *mats, DevStatus *st, int aux, bool write_vc,
                  bool general, int64_t n, int64_t max_blocks, hipStream_t s);
// Lattice scan: the new partition from the destination marks and particle counts.
// prev_used: tile_used of the partition the particles move from
void launch_lattice_scan(Lattice L, Partition Pn, Binning B, const uint32_t *prev_used, int64_t n, DevStatus *st,
                         hipStream_t s);
// G2P update: in place (scatter == false), or written into the next stable block order.
// aux: kAux* channels the scatter carries; nonaffine: the FLIP transfer (velocity
// gradient from the weight gradients, C kept).
void launch_g2p_update(bool scatter, State A, State B, Partition P, Partition Pn, Lattice L, const float *grid,
                       Binning Bn, int64_t max_blocks, hipStream_t s, int wg = 256, const float *gstar = nullptr,
                       int aux = 0, bool nonaffine = false);
// Affine FLIP transfers: P2G of the momentum (into tiles) and of the stress
// (into tiles2), and their grid stages (grid: m, v; gstar: m, vstar).
void launch_p2g_aflip(State A, Partition P, Lattice L, const DevMaterial *mats, const float *sig, int64_t stride,
                      float *tiles, float *tiles2, DevStatus *st, int64_t max_blocks, hipStream_t s);
// FLIP transfer: P2G of W m v (into tiles) and of W (-dt vol P F^T grad W) (into tiles2).
void launch_p2g_flip(State A, Partition P, Lattice L, const DevMaterial *mats, const float *sig, int64_t stride,
                     float *tiles, float *tiles2, DevStatus *st, int64_t max_blocks, hipStream_t s);
void launch_grid_aflip(int mode, Partition P, Lattice L, const float *tiles, const float *tiles2, float *grid,
                       float *gstar, float *gvmax, int64_t max_blocks, hipStream_t s, uint32_t *vclear = nullptr,
                       int32_t nclear = 0);
// view: a const zmpm_particles* (device pointers inside); its particle i becomes resident
// particle off + i with caller index off + i (one object of a list of particle objects)
void launch_import(State dst, int64_t cap, const void *view, int64_t off, const DevMaterial *mats, int32_t nmat,
                   DevStatus *st, hipStream_t s);
// aux: the kAux* channels the resident state keeps (the others are exported as their initial values);
// the resident particles of caller index off .. off + view->n - 1 go to the view's entries 0 .. n - 1
void launch_export(State src, int64_t n_res, const void *view, int64_t off, const DevMaterial *mats, int aux,
                   hipStream_t s);
// Particle volume dx^3 / (particles in the cell) and mass rho vol into caller buffers
// (one channel each, ZMPM_LAYOUT_*) at the caller particle index; cnt: scratch of 64
// words per grid block.
// Level-set colliders on the grid's updated velocities (vmax: fold the final
// speeds into the per-block maxima).
void launch_colliders(Partition P, Lattice L, float *grid, float *gvmax, int64_t max_blocks, bool vmax,
                      hipStream_t s);
// Advance every collider by its scene's dt (dt < 0) or by dt (StepZSBoundary).
void launch_step_colliders(Lattice L, float dt, hipStream_t s);

// Boundary particles on the grid (ZSBoundaryPrimitiveToZSGrid + ApplyGridBoundaryOnZSGrid):
// acc holds 7 int64 channels (m, m v, m nrm) x 64 cells per grid block, scale_bits one word.
struct BoundaryParticles {
    int64_t n;
    const float *x, *v, *m, *nrm;
    int32_t scene, type;
    unsigned long long *acc;
    uint32_t *scale_bits;  // max over particles of |m| max(1, |v|) as float bits
};
constexpr int kBndChannels = 7;
// apply false: only the sums (for ComputeParticleBeta's boundary grid)
void launch_boundary_particles(BoundaryParticles bp, Partition P, Lattice L, float *grid, float *gvmax,
                               int64_t max_gblocks, hipStream_t s, bool apply = true);
// Return mapping of the plastic particles' F (ZSReturnMapping; Cam-clay also updates logJp).
void launch_return_mapping(State A, int64_t n, const DevMaterial *mats, hipStream_t s);
// Reset the kAux* channels of mask: beta = the material's flip_beta, logJp = kLogJpInit.
void launch_init_aux(State A, int64_t n, const DevMaterial *mats, int mask, hipStream_t s);
// ComputeParticleBeta with the context's colliders and boundary particles (bp.n > 0: their sums on
// the partition P); dt < 0: each scene's dt.
void launch_particle_beta(State A, int64_t n, Lattice L, const DevMaterial *mats, float dt, float jp_critical,
                          BoundaryParticles bp, Partition P, hipStream_t s);

// Morton order of a caller's particle buffer (ZMPM_LAYOUT_*, nch channels, n particles):
// codes of the positions at channels x_chn..x_chn+2 in the box lo .. lo + ext, and the
// permutation of every channel of src into dst by order (dst[k] = src[order[k]]).
void launch_morton_codes(int32_t layout, int32_t nch, int64_t n, const float *src, int32_t x_chn, const float lo[3],
                         const float ext[3], uint32_t *codes, hipStream_t s);
void launch_permute_channels(int32_t layout, int32_t nch, int64_t n, const float *src, float *dst,
                             const uint32_t *order, hipStream_t s);

struct VolumeOut {
    int32_t layout, tile_channels;
    float *vol;
    int32_t vol_chn;
    float *mass;
    int32_t mass_chn;
};
void launch_particle_volume(State A, int64_t n, Partition P, Lattice L, const DevMaterial *mats, uint32_t *cnt,
                            int64_t max_gblocks, VolumeOut out, DevStatus *st, hipStream_t s);

}  // namespace zmpm
