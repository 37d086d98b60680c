// Host side of libzmpm: context, resident state, node-granular stages and the
// fused substep chain, behind the C ABI of include/zmpm.h.
//
// Every entry point catches everything and turns it into a status code plus a
// thread-local message (the reference's C-API does the same with
// LastError::catched, zeno/include/zeno/extra/CAPI.h:133-163).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/zmpm.h"
#include "zmpm_internal.h"

using namespace zmpm;

namespace {

thread_local std::string g_last_error;

struct ZmpmError : std::runtime_error {
    int32_t code;
    ZmpmError(int32_t c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(expr)                                                                                  \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            throw ZmpmError(ZMPM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

template <typename F> int32_t guarded(F &&f) {
    try {
        f();
        return ZMPM_OK;
    } catch (const ZmpmError &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return ZMPM_ERR_ARG;
    } catch (...) {
        g_last_error = "unknown error";
        return ZMPM_ERR_ARG;
    }
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

struct zmpm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t cap = 0;
    Lattice L{};                           // L.scenes points at scenes_dev
    std::vector<SceneParams> scenes_host;  // as created (dx, domain, default step parameters)
    SceneParams *scenes_dev = nullptr;
    uint32_t *vmax = nullptr;              // per scene: max |v|^2 (float bits), reduced on request
    float *gvmax = nullptr;                // per grid block: max |v|^2 of the latest grid update
    int key_bits = 0;
    int64_t n = 0;
    bool partitioned = false;  // the partition describes the current particle positions
    bool loaded_sorted = false;  // the resident state was loaded from a save (saves are block-sorted)
    int wg_move = 256, wg_update = 256;  // G2P workgroup sizes (zmpm_autotune)
    bool p2g_sparse = true;              // one-wave P2G for blocks of <= 64 particles (zmpm_autotune)
    bool fuse = false;                   // APIC: the fused G2P / P2G pass (zmpm_autotune, zmpm_set_fused);
                                         // while on, every P2G of the context uses its chunk-flush grouping
    bool tiles_valid = false;            // fused pass: tiles hold the P2G of the current state and partition
    uint32_t *arr = nullptr;             // fused pass: arriving particles' old indices at their new slots
    bool fused_path() const { return fuse && transfer == ZMPM_TRANSFER_APIC; }
    bool has_plastic = false;            // a material with plasticity: return mapping after G2P
    bool has_stvk = false;               // a solid model other than fixed corotated (no adjoint)
    bool has_aniso = false;              // an ARAP fiber term: the particles' fiber is kept (no adjoint)
    // the stress kernel with every F-based model, or its fixed-corotated-only form
    bool stress_general() const { return has_stvk || has_aniso; }
    bool pp_fiber = false;               // the particles were imported with their own fiber (kept from then on)
    bool cols_move = false;              // a collider with a translation or scale rate
    bool has_nacc = false;               // a Cam-clay material: the particles' logJp is kept
    bool beta_kept = false;              // zmpm_compute_particle_beta wrote the particles' beta
    BoundaryParticles bnd{};             // zmpm_set_boundary_particles (bnd.n == 0: none)
    bool pp_lame = false;                // the particles were imported with their own mu, lam
    bool pp_mass = false;                // the particles carry their own m, vol (imported, or particle volume)
    bool grid_ready = false;   // the grid belongs to the current partition

    State st[2]{};
    int cur = 0;
    Partition part[2]{};
    int pcur = 0;
    int grid_part = 0;  // the partition the grid values belong to

    uint32_t *keys = nullptr, *skeys = nullptr, *perm = nullptr;
    void *sort_scratch = nullptr;
    Binning bin{};
    float *tiles = nullptr, *grid = nullptr;
    int64_t tiles_blocks = 0;              // parts the tile buffer holds
    int64_t parts_cap = 0;                 // entries of the part-indexed arrays
    uint32_t psize[2] = {kNoSplit, kNoSplit};  // part shift each partition buffer was cut with
    // the fused chain works per block (one part each); everything else splits large blocks
    uint32_t wanted_part_shift() const { return fused_path() ? kNoSplit : kPartShift; }
    int64_t grid_blocks = 0;               // grid blocks the grid buffer holds
    int transfer = ZMPM_TRANSFER_APIC;     // ZMPM_TRANSFER_*
    float *tiles2 = nullptr, *gstar = nullptr;  // FLIP transfers: stress tiles and the vstar grid
    bool flip_style() const { return transfer != ZMPM_TRANSFER_APIC; }
    // the particle channels the G2P scatter carries: beta while a FLIP transfer reads it (or
    // after zmpm_compute_particle_beta), logJp while a Cam-clay material updates it
    int aux() const {
        return ((flip_style() || beta_kept) ? kAuxBeta : 0) | (has_nacc ? kAuxLogJp : 0) | (pp_lame ? kAuxLame : 0) |
               ((has_aniso || pp_fiber) ? kAuxFiber : 0) | (pp_mass ? kAuxMass : 0);
    }
    // the stress reads the particles' mu, lam unless the transfer is FLIP (its P2G has no override);
    // P2G and the stress read the particles' m, vol under every transfer (Transfer.cu:857-858, 905-906)
    void sync_pp_flag() {
        L.pp_lame = (pp_lame && transfer != ZMPM_TRANSFER_FLIP) ? 1 : 0;
        L.pp_mass = pp_mass ? 1 : 0;
    }
    float *sig = nullptr;  // stress terms of the fixed-corotated particles, planes of cap entries
    DevStatus *status = nullptr;
    DevMaterial *mats = nullptr;
    int32_t nmat = 0;
    std::vector<zmpm_material> host_mats;
    std::vector<void *> allocations;

    // differentiation scratch, allocated on first use
    float *grid_v = nullptr;   // updated grid velocities (the combined momentum stays in grid)
    float *gbar = nullptr;     // grid adjoint
    float *pay = nullptr;      // G2P adjoint payload: a (3, cap), B (9, cap)

    bool graph_fused = false;             // the captured chain is the fused one
    uint32_t graph_shift = kNoSplit;      // the part shift the captured chain cuts with
    hipGraph_t graph = nullptr;           // zmpm_graph_capture
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t capture_stream = nullptr;

    uint32_t *host_counts = nullptr;  // pinned, device-mapped: block counts of the latest partition
    // Launch sizes: the latest block counts the device has published, with headroom
    // (kernels loop over blocks beyond the grid, so a stale estimate stays correct).
    int64_t max_pblocks() const {
        const int64_t hi = std::min<int64_t>(L.size, std::max<int64_t>(n, 1));
        const int64_t est = host_counts ? (int64_t)((volatile uint32_t *)host_counts)[0] : 0;
        return est > 0 ? std::min<int64_t>(hi, est + est / 4 + 64) : hi;
    }
    // work units of P2G and the G2P passes: particle blocks plus extra parts, same headroom
    int64_t max_units() const {
        const int64_t est = host_counts ? (int64_t)((volatile uint32_t *)host_counts)[2] : 0;
        return std::min<int64_t>(parts_cap, max_pblocks() + (est > 0 ? est + est / 4 + 64 : 0));
    }
    int64_t max_gblocks() const {
        const int64_t hi = std::min<int64_t>(L.size, 8 * std::max<int64_t>(n, 1));
        const int64_t est = host_counts ? (int64_t)((volatile uint32_t *)host_counts)[1] : 0;
        return est > 0 ? std::min<int64_t>(hi, est + est / 4 + 64) : hi;
    }

    void *alloc(size_t bytes) {
        void *p = nullptr;
        HIP_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 256)));
        allocations.push_back(p);
        return p;
    }
    ~zmpm_ctx() {
        if (device >= 0) (void)hipSetDevice(device);
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (capture_stream) (void)hipStreamDestroy(capture_stream);
        for (void *p : allocations) (void)hipFree(p);
        if (host_counts) (void)hipHostFree(host_counts);
    }
    void set_device() { HIP_CHECK(hipSetDevice(device)); }
};

namespace {

StepArgs step_args(float dt, const float accel[3], int32_t boundary, int32_t pad) {
    StepArgs a{};
    a.dt = dt;
    for (int k = 0; k < 3; ++k) a.accel[k] = accel ? accel[k] : 0.0f;
    a.boundary = boundary;
    a.pad = pad;
    return a;
}

// Per-scene max |v|^2 of the latest grid update into ctx->vmax (stream ordered).
void reduce_vmax(zmpm_ctx *c) {
    fill_u32(c->vmax, 0u, c->L.nscenes, c->stream);
    launch_reduce_vmax(c->gvmax, c->part[c->grid_part], c->max_gblocks(), c->vmax, c->stream);
}

// Write step parameters into every scene's entry: the device table (stream
// ordered) and the host mirror whose entry 0 travels by value in kernel arguments.
void set_steps(zmpm_ctx *c, const StepArgs &a, int mask) {
    for (SceneParams &sc : c->scenes_host) {
        // the fused pass's tiles hold the next P2G, whose stress terms scale with dt
        if ((mask & SET_DT) && sc.dt != a.dt) c->tiles_valid = false;
        if (mask & SET_DT) sc.dt = a.dt;
        if (mask & SET_ACCEL)
            for (int k = 0; k < 3; ++k) sc.accel[k] = a.accel[k];
        if (mask & SET_BOUNDARY) {
            sc.boundary = a.boundary;
            sc.pad = a.pad;
        }
    }
    c->L.s0 = c->scenes_host[0];
    launch_set_steps(c->scenes_dev, c->L.nscenes, a, mask, c->stream);
}

void check_ctx(const zmpm_ctx *c) {
    if (!c) throw ZmpmError(ZMPM_ERR_ARG, "null context");
}

// Device status words -> error codes (synchronising).
void check_status(zmpm_ctx *c) {
    DevStatus h{};
    HIP_CHECK(hipMemcpyAsync(&h, c->status, sizeof(DevStatus), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    if (h.bad_material) throw ZmpmError(ZMPM_ERR_ARG, "particle material index outside the material table");
    if (h.domain_error)
        throw ZmpmError(ZMPM_ERR_DOMAIN, "a particle left the grid lattice or has a non-finite position");
    if (h.cfl_error)
        throw ZmpmError(ZMPM_ERR_DOMAIN, "a particle moved more than one 4^3 block in one substep (dt above the CFL limit)");
    if (h.bin_mismatch || h.scan_stall)
        throw ZmpmError(ZMPM_ERR_STATE, "internal binning inconsistency (bin mismatch or look-back stall)");
}

// Bin the resident particles: keys -> stable sort -> permute the state into
// sorted order -> particle blocks and grid blocks (B + {0,1}^3).
// part_shift: the parts (of 2^part_shift particles) to cut the blocks into (0: what the context's
// chain wants).
void full_partition(zmpm_ctx *c, uint32_t part_shift = 0) {
    hipStream_t s = c->stream;
    c->L.part_shift = part_shift ? part_shift : c->wanted_part_shift();
    c->psize[c->pcur] = c->L.part_shift;
    State A = c->st[c->cur], B = c->st[c->cur ^ 1];
    // a saved state is already in stable block order: the sort would be the identity
    // (a state that is not is reported as a binning inconsistency at the next sync)
    launch_bin_keys(A.ch, -1, c->n, c->mats, c->L, c->keys, c->status, s, c->loaded_sorted);
    if (c->loaded_sorted) {
        launch_partition_from_keys(c->keys, c->n, c->L, c->part[c->pcur], c->bin, A, c->mats, c->status, s);
        c->loaded_sorted = false;
    } else {
        // stable sort by block (key bits above the 6 cell bits)
        radix_sort_pairs(c->keys, nullptr, c->n, c->L.key_shift, c->key_bits, c->skeys, c->perm, nullptr,
                         c->sort_scratch, s);
        launch_gather_state(A, c->cap, B, c->cap, c->perm, c->n, s);
        c->cur ^= 1;
        launch_partition_from_keys(c->skeys, c->n, c->L, c->part[c->pcur], c->bin, c->st[c->cur], c->mats,
                                   c->status, s);
    }
    HIP_CHECK(hipGetLastError());
    c->partitioned = true;
    c->grid_ready = false;
    c->grid_part = c->pcur;
    c->tiles_valid = false;
}

}  // namespace

extern "C" {

const char *zmpm_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"

namespace {

zmpm_ctx *create_context(const zmpm_scene *scenes, int32_t nscenes, int64_t capacity, int32_t device) {
    if (!scenes || nscenes <= 0 || nscenes > kMaxScenes)
        throw ZmpmError(ZMPM_ERR_ARG, "scene count must be in [1, 32768]");
    if (capacity <= 0) throw ZmpmError(ZMPM_ERR_ARG, "capacity must be positive");
    // resident tiles are addressed with 32-bit float offsets: kTileCh (35) channels x capacity < 2^32
    if (state_floats(capacity) > (int64_t)0xffffffffll)
        throw ZmpmError(ZMPM_ERR_ARG, "capacity above 122,713,344 particles");
    std::vector<SceneParams> sp(nscenes);
    int64_t lat = 0;
    for (int s = 0; s < nscenes; ++s) {
        const zmpm_scene &z = scenes[s];
        if (!(z.dx > 0.0f) || z.n_cells <= 0) throw ZmpmError(ZMPM_ERR_ARG, "dx and n_cells must be positive");
        if (z.boundary != ZMPM_BOUNDARY_NONE && z.boundary != ZMPM_BOUNDARY_STICKY_BOX)
            throw ZmpmError(ZMPM_ERR_ARG, "unknown boundary type");
        SceneParams &c = sp[s];
        c.dx = z.dx;
        c.dxinv = 1.0f / z.dx;
        c.Dinv = 4.0f * c.dxinv * c.dxinv;  // quadratic B-spline D^-1 (Transfer.cu:854)
        c.dt = z.dt;
        for (int k = 0; k < 3; ++k) c.accel[k] = z.gravity[k];
        c.boundary = z.boundary;
        c.pad = z.pad;
        c.n_cells = z.n_cells;
        c.nb = (z.n_cells + 3) / 4 + 3;  // blocks -1 .. ceil(n/4)+1: stencils reach one block past the domain
        c.lat_off = (int32_t)lat;
        lat += (int64_t)c.nb * c.nb * c.nb;
        // 32-bit lattice indices (binning keys carry the 6 cell bits only below 2^26 entries)
        if (lat >= ((int64_t)1 << 31)) throw ZmpmError(ZMPM_ERR_ARG, "lattice of 2^31 blocks or more");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0)
        throw ZmpmError(ZMPM_ERR_NODEVICE, "no HIP device with the requested ordinal");
    auto *c = new zmpm_ctx();
    try {
        c->device = device;
        c->set_device();
        hipDeviceProp_t prop;
        HIP_CHECK(hipGetDeviceProperties(&prop, c->device));
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            throw ZmpmError(ZMPM_ERR_NODEVICE, std::string("libzmpm is built for gfx950, device is ") +
                                                   prop.gcnArchName);
        c->cap = capacity;
        c->scenes_host = sp;
        c->scenes_dev = (SceneParams *)c->alloc(sizeof(SceneParams) * nscenes);
        HIP_CHECK(hipMemcpy(c->scenes_dev, sp.data(), sizeof(SceneParams) * nscenes, hipMemcpyHostToDevice));
        c->vmax = (uint32_t *)c->alloc(sizeof(uint32_t) * nscenes);
        HIP_CHECK(hipMemset(c->vmax, 0, sizeof(uint32_t) * nscenes));
        c->L.origin = -1;
        c->L.nscenes = nscenes;
        c->L.size = lat;
        c->L.scenes = c->scenes_dev;
        c->L.s0 = sp[0];
        int bits = 0;
        while (((int64_t)1 << bits) < c->L.size) ++bits;
        c->L.key_shift = bits + 6 <= 32 ? 6 : 0;
        c->key_bits = bits + c->L.key_shift;
        const int64_t cap = c->cap, L = c->L.size;
        // per-block arrays by the most blocks there can be, not by the lattice
        const int64_t max_pb = std::min<int64_t>(L, cap);
        // a block of d particles has ceil(d / kPartParticles) - 1 extra parts
        const int64_t parts_cap = max_pb + cap / kPartParticles + 1;
        c->parts_cap = parts_cap;
        c->L.part_shift = kPartShift;
        c->grid_blocks = std::min<int64_t>(L, 8 * max_pb);  // grid blocks are B + {0,1}^3: <= 8 per particle block
        c->gvmax = (float *)c->alloc(sizeof(float) * c->grid_blocks);
        HIP_CHECK(hipMemset(c->gvmax, 0, sizeof(float) * c->grid_blocks));
        for (int k = 0; k < 2; ++k) {
            c->st[k].ch = (float *)c->alloc(sizeof(float) * state_floats(cap));
            Partition &P = c->part[k];
            P.lat_pid = (int32_t *)c->alloc(sizeof(int32_t) * L);
            P.lat_gid = (int32_t *)c->alloc(sizeof(int32_t) * L);
            P.pblk_lat = (uint32_t *)c->alloc(sizeof(uint32_t) * max_pb);
            P.pblk_start = (uint32_t *)c->alloc(sizeof(uint32_t) * (max_pb + 1));
            P.gblk_lat = (uint32_t *)c->alloc(sizeof(uint32_t) * c->grid_blocks);
            P.pblk_scene = (uint32_t *)c->alloc(sizeof(uint32_t) * max_pb);
            P.gblk_scene = (uint32_t *)c->alloc(sizeof(uint32_t) * c->grid_blocks);
            P.counts = (uint32_t *)c->alloc(sizeof(uint32_t) * 4);
            P.pblk_x = (uint32_t *)c->alloc(sizeof(uint32_t) * (max_pb + 1));
            P.lat_xf = (uint8_t *)c->alloc((size_t)L);
            HIP_CHECK(hipMemset(P.lat_xf, 0, (size_t)L));
            P.part_blk = (uint32_t *)c->alloc(sizeof(uint32_t) * parts_cap);
            P.part_lo = (uint32_t *)c->alloc(sizeof(uint32_t) * parts_cap);
            P.parts_cap = (uint32_t)parts_cap;
            // every tile is visited by the first scan into this buffer (its entries start undefined)
            P.tile_used = (uint32_t *)c->alloc(sizeof(uint32_t) * ((L + kScanTile - 1) / kScanTile));
            fill_u32(P.tile_used, 1u, (L + kScanTile - 1) / kScanTile, nullptr);
            HIP_CHECK(hipMemset(P.counts, 0, sizeof(uint32_t) * 4));
        }
        Binning &B = c->bin;
        B.lat_flag = (uint32_t *)c->alloc(sizeof(uint32_t) * L);
        B.dcount = (uint32_t *)c->alloc(sizeof(uint32_t) * L);
        B.cnt27 = (uint32_t *)c->alloc(sizeof(uint32_t) * kDirs * parts_cap);
        B.code = (uint32_t *)c->alloc(sizeof(uint32_t) * cap);
        B.mvlist = (uint32_t *)c->alloc(sizeof(uint32_t) * cap);
        c->arr = (uint32_t *)c->alloc(sizeof(uint32_t) * cap);
        B.scan_tiles = (L + kScanTile - 1) / kScanTile;
        B.tile_mark = (uint32_t *)c->alloc(sizeof(uint32_t) * B.scan_tiles);
        int64_t nbmax = 0;
        for (const SceneParams &q : sp) nbmax = std::max<int64_t>(nbmax, q.nb);
        B.mark_radius = (int32_t)((2 * nbmax * nbmax + 2 * nbmax + 2 + kScanTile - 1) / kScanTile + 1);
        B.vmax_out = c->vmax;
        HIP_CHECK(hipMemset(B.tile_mark, 0, sizeof(uint32_t) * B.scan_tiles));
        B.scan_p.status = (uint64_t *)c->alloc(sizeof(uint64_t) * B.scan_tiles);
        B.scan_g.status = (uint64_t *)c->alloc(sizeof(uint64_t) * B.scan_tiles);
        B.scan_p.tile_ctr = (uint32_t *)c->alloc(sizeof(uint32_t) * 4);
        B.scan_g.tile_ctr = B.scan_p.tile_ctr;
        B.scan_p.stall_flag = B.scan_p.tile_ctr + 1;
        B.scan_g.stall_flag = B.scan_p.tile_ctr + 1;
        HIP_CHECK(hipMemset(B.lat_flag, 0, sizeof(uint32_t) * L));
        HIP_CHECK(hipMemset(B.dcount, 0, sizeof(uint32_t) * L));
        HIP_CHECK(hipMemset(B.scan_p.tile_ctr, 0, sizeof(uint32_t) * 4));
        HIP_CHECK(hipHostMalloc((void **)&c->host_counts, 4 * sizeof(uint32_t), hipHostMallocMapped));
        std::memset(c->host_counts, 0, 4 * sizeof(uint32_t));
        HIP_CHECK(hipHostGetDevicePointer((void **)&B.host_counts, c->host_counts, 0));
        c->keys = (uint32_t *)c->alloc(sizeof(uint32_t) * cap);
        c->skeys = (uint32_t *)c->alloc(sizeof(uint32_t) * cap);
        c->perm = (uint32_t *)c->alloc(sizeof(uint32_t) * cap);
        c->sort_scratch = c->alloc(radix_scratch_bytes(cap));
        c->tiles = (float *)c->alloc(sizeof(float) * 4 * kTileNodes * parts_cap);
        c->tiles_blocks = parts_cap;
        c->sig = (float *)c->alloc(sizeof(float) * 9 * cap);
        c->grid = (float *)c->alloc(sizeof(float) * 256 * c->grid_blocks);
        c->status = (DevStatus *)c->alloc(sizeof(DevStatus));
        HIP_CHECK(hipMemset(c->status, 0, sizeof(DevStatus)));
        c->mats = (DevMaterial *)c->alloc(sizeof(DevMaterial) * kMaxMaterials);
        HIP_CHECK(hipDeviceSynchronize());
    } catch (...) {
        delete c;
        throw;
    }
    return c;
}

}  // namespace

extern "C" {

int32_t zmpm_create(zmpm_ctx **out, const zmpm_config *cfg) {
    return guarded([&] {
        if (!out || !cfg) throw ZmpmError(ZMPM_ERR_ARG, "null argument");
        *out = nullptr;
        // one scene; step parameters come with each stage call or zmpm_substeps
        zmpm_scene z{};
        z.dx = cfg->dx;
        z.n_cells = cfg->n_cells;
        z.gravity[1] = -9.8f;  // UpdateZSGrid default gravity (MPMPipeline.cu:947)
        z.boundary = ZMPM_BOUNDARY_NONE;
        *out = create_context(&z, 1, cfg->capacity, cfg->device);
    });
}

int32_t zmpm_create_scenes(zmpm_ctx **out, const zmpm_scene *scenes, int32_t nscenes, int64_t capacity,
                           int32_t device) {
    return guarded([&] {
        if (!out) throw ZmpmError(ZMPM_ERR_ARG, "null argument");
        *out = nullptr;
        *out = create_context(scenes, nscenes, capacity, device);
    });
}

int32_t zmpm_num_scenes(const zmpm_ctx *ctx) { return ctx ? ctx->L.nscenes : 0; }

int32_t zmpm_destroy(zmpm_ctx *ctx) {
    return guarded([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        delete ctx;
    });
}

int32_t zmpm_set_stream(zmpm_ctx *ctx, void *stream) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->stream = (hipStream_t)stream;
    });
}

int32_t zmpm_sync(zmpm_ctx *ctx) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->set_device();
        check_status(ctx);
    });
}

int32_t zmpm_set_materials(zmpm_ctx *ctx, const zmpm_material *mats, int32_t count) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->tiles_valid = false;  // the fused pass's tiles no longer match the state
        if (!mats || count <= 0 || count > kMaxMaterials)
            throw ZmpmError(ZMPM_ERR_ARG, "material count must be in [1, 4096]");
        std::vector<DevMaterial> dm(count);
        for (int i = 0; i < count; ++i) {
            const zmpm_material &m = mats[i];
            if (m.model != ZMPM_MODEL_FCR && m.model != ZMPM_MODEL_JFLUID && m.model != ZMPM_MODEL_STVK_HENCKY &&
                m.model != ZMPM_MODEL_NEO_HOOKEAN && m.model != ZMPM_MODEL_STABLE_NEO_HOOKEAN)
                throw ZmpmError(ZMPM_ERR_ARG, "unknown material model");
            if (m.plasticity != ZMPM_PLASTIC_NONE && m.plasticity != ZMPM_PLASTIC_DRUCKER_PRAGER &&
                m.plasticity != ZMPM_PLASTIC_VON_MISES && m.plasticity != ZMPM_PLASTIC_CAM_CLAY)
                throw ZmpmError(ZMPM_ERR_ARG, "unknown plasticity model");
            if (m.plasticity != ZMPM_PLASTIC_NONE && m.model != ZMPM_MODEL_STVK_HENCKY)
                throw ZmpmError(ZMPM_ERR_ARG, "plasticity needs the StVK-Hencky model (as ConfigConstitutiveModel)");
            if (m.plasticity == ZMPM_PLASTIC_DRUCKER_PRAGER && !(m.friction_angle >= 0.0f && m.friction_angle < 90.0f))
                throw ZmpmError(ZMPM_ERR_ARG, "friction angle must be in [0, 90) degrees");
            if (m.plasticity == ZMPM_PLASTIC_VON_MISES && !(m.yield_stress >= 0.0f))
                throw ZmpmError(ZMPM_ERR_ARG, "yield stress must be >= 0");
            if (m.plasticity == ZMPM_PLASTIC_CAM_CLAY &&
                !(m.friction_angle >= 0.0f && m.friction_angle < 90.0f && m.nacc_beta >= 0.0f && m.nacc_xi >= 0.0f))
                throw ZmpmError(ZMPM_ERR_ARG, "Cam-clay needs friction angle in [0, 90), nacc_beta >= 0, nacc_xi >= 0");
            if (m.aniso != ZMPM_ANISO_NONE && m.aniso != ZMPM_ANISO_ARAP)
                throw ZmpmError(ZMPM_ERR_ARG, "unknown anisotropic model");
            if (m.aniso == ZMPM_ANISO_ARAP && (m.model == ZMPM_MODEL_JFLUID || !(m.aniso_strength >= 0.0f) ||
                                               !std::isfinite(m.aniso_strength)))
                throw ZmpmError(ZMPM_ERR_ARG, "the ARAP fiber term needs an F-based model and a finite strength >= 0");
            if (!(m.vol > 0.0f) || !(m.rho > 0.0f)) throw ZmpmError(ZMPM_ERR_ARG, "material vol and rho must be > 0");
            if (m.scene < 0 || m.scene >= ctx->L.nscenes) throw ZmpmError(ZMPM_ERR_ARG, "material scene out of range");
            DevMaterial d{};
            d.model = m.model;
            // Lame parameters as FixedCorotated<float>{E, nu} (Generation.cu:41)
            d.mu = m.E / (2.0f * (1.0f + m.nu));
            d.lam = m.E * m.nu / ((1.0f + m.nu) * (1.0f - 2.0f * m.nu));
            d.mass = m.rho * m.vol;  // m = vol * density (Generation.cpp:944-946)
            d.vol = m.vol;
            d.bulk = m.bulk;
            d.scene = m.scene;
            d.rho = m.rho;
            d.plasticity = m.plasticity;
            // alpha = sqrt(2/3) 2 sin(phi) / (3 - sin(phi)) (Klar et al. 2016, eq. 27, no cohesion)
            const double sphi = std::sin((double)m.friction_angle * 3.14159265358979323846 / 180.0);
            d.dp_alpha = (float)(std::sqrt(2.0 / 3.0) * 2.0 * sphi / (3.0 - sphi));
            d.yield_stress = m.yield_stress;
            d.flip_beta = m.flip_beta;
            // Cam-clay: critical-state slope M = 6 sin(phi) / (3 - sin(phi)) (the Mohr-Coulomb cone
            // in triaxial compression, 3D), bulk modulus K = 2 mu / 3 + lam of tau = 2 mu e + lam tr(e) I
            d.nacc_M = (float)(6.0 * sphi / (3.0 - sphi));
            d.nacc_beta = m.nacc_beta;
            d.nacc_xi = m.nacc_xi;
            d.nacc_bulk = 2.0f * d.mu / 3.0f + d.lam;
            // AnisotropicArap{E, nu, strength} (Generation.cu:63): the material's mu, not a particle's
            d.aniso_mu = m.aniso == ZMPM_ANISO_ARAP ? m.aniso_strength * d.mu : 0.0f;
            dm[i] = d;
        }
        bool plastic = false, stvk = false, nacc = false, aniso = false;
        for (int i = 0; i < count; ++i) {
            plastic = plastic || mats[i].plasticity != ZMPM_PLASTIC_NONE;
            stvk = stvk || (mats[i].model != ZMPM_MODEL_FCR && mats[i].model != ZMPM_MODEL_JFLUID);
            nacc = nacc || mats[i].plasticity == ZMPM_PLASTIC_CAM_CLAY;
            aniso = aniso || mats[i].aniso == ZMPM_ANISO_ARAP;
        }
        ctx->set_device();
        HIP_CHECK(hipMemcpyAsync(ctx->mats, dm.data(), sizeof(DevMaterial) * count, hipMemcpyHostToDevice,
                                 ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        ctx->host_mats.assign(mats, mats + count);
        ctx->nmat = count;
        ctx->has_plastic = plastic;
        ctx->has_stvk = stvk;
        // resident particles whose logJp was not kept start from the reference's -0.04
        if (nacc && !ctx->has_nacc && ctx->n > 0) {
            launch_init_aux(ctx->st[ctx->cur], ctx->n, ctx->mats, kAuxLogJp, ctx->stream);
            HIP_CHECK(hipGetLastError());
        }
        ctx->has_nacc = nacc;
        // resident particles whose fiber was not kept: the import default (0, 1, 0)
        if (aniso && !ctx->has_aniso && !ctx->pp_fiber && ctx->n > 0) {
            launch_init_aux(ctx->st[ctx->cur], ctx->n, ctx->mats, kAuxFiber, ctx->stream);
            HIP_CHECK(hipGetLastError());
        }
        ctx->has_aniso = aniso;
    });
}

namespace {

// Host-side validation of one view: layout, count, and for AoSoA32 every present
// attribute's channels inside the tile (a channel past the tile would make the
// import read, and the export write, memory the caller does not own).
void check_view(const zmpm_particles *v) {
    if (!v) throw ZmpmError(ZMPM_ERR_ARG, "null particle view");
    if (v->n < 0) throw ZmpmError(ZMPM_ERR_ARG, "negative particle count");
    if (v->n > 0 && !v->x) throw ZmpmError(ZMPM_ERR_ARG, "particle view without positions");
    if (v->layout < ZMPM_LAYOUT_SOA || v->layout > ZMPM_LAYOUT_AOSOA32)
        throw ZmpmError(ZMPM_ERR_ARG, "unknown particle layout");
    if ((v->mu == nullptr) != (v->lam == nullptr)) throw ZmpmError(ZMPM_ERR_ARG, "mu and lam go together");
    if ((v->m == nullptr) != (v->vol == nullptr)) throw ZmpmError(ZMPM_ERR_ARG, "m and vol go together");
    if (v->layout != ZMPM_LAYOUT_AOSOA32) return;
    if (v->tile_channels <= 0) throw ZmpmError(ZMPM_ERR_ARG, "AoSoA32 layout needs tile_channels > 0");
    struct Field {
        const void *p;
        int32_t chn, width;
        const char *name;
    };
    const Field fields[] = {{v->x, v->x_chn, 3, "x"},         {v->v, v->v_chn, 3, "v"},
                            {v->C, v->C_chn, 9, "C"},         {v->F, v->F_chn, 9, "F"},
                            {v->J, v->J_chn, 1, "J"},         {v->beta, v->beta_chn, 1, "beta"},
                            {v->logJp, v->logJp_chn, 1, "logJp"}, {v->mu, v->mu_chn, 1, "mu"},
                            {v->lam, v->lam_chn, 1, "lam"},   {v->a, v->a_chn, 3, "a"},
                            {v->m, v->m_chn, 1, "m"},         {v->vol, v->vol_chn, 1, "vol"}};
    for (const Field &f : fields)
        if (f.p && (f.chn < 0 || (int64_t)f.chn + f.width > v->tile_channels))
            throw ZmpmError(ZMPM_ERR_ARG, std::string("channels of \"") + f.name + "\" outside the tile (chn " +
                                              std::to_string(f.chn) + ", width " + std::to_string(f.width) +
                                              ", tile_channels " + std::to_string(v->tile_channels) + ")");
}

int64_t check_views(const zmpm_particles *views, int32_t nviews) {
    if (nviews < 1 || !views) throw ZmpmError(ZMPM_ERR_ARG, "at least one particle view is needed");
    int64_t total = 0;
    for (int k = 0; k < nviews; ++k) {
        check_view(&views[k]);
        total += views[k].n;
    }
    return total;
}

void import_views(zmpm_ctx *ctx, const zmpm_particles *views, int32_t nviews) {
    check_ctx(ctx);
    ctx->tiles_valid = false;  // the fused pass's tiles no longer match the state
    const int64_t total = check_views(views, nviews);
    if (total > ctx->cap) throw ZmpmError(ZMPM_ERR_CAPACITY, "more particles than the context capacity");
    if (ctx->nmat <= 0) throw ZmpmError(ZMPM_ERR_STATE, "zmpm_set_materials must precede zmpm_import");
    // an optional attribute is per particle when every object carries it
    bool lame = true, fiber = true, mass = true;
    for (int k = 0; k < nviews; ++k) {
        lame = lame && views[k].mu;
        fiber = fiber && views[k].a;
        mass = mass && views[k].m;
    }
    ctx->set_device();
    ctx->n = total;
    ctx->pp_lame = lame;
    ctx->pp_fiber = fiber;
    ctx->pp_mass = mass;
    ctx->sync_pp_flag();
    fill_u32((uint32_t *)ctx->status, 0u, sizeof(DevStatus) / 4, ctx->stream);
    int64_t off = 0;
    for (int k = 0; k < nviews; ++k) {
        launch_import(ctx->st[ctx->cur], ctx->cap, &views[k], off, ctx->mats, ctx->nmat, ctx->status, ctx->stream);
        off += views[k].n;
    }
    HIP_CHECK(hipGetLastError());
    full_partition(ctx);
    check_status(ctx);
}

void export_views(zmpm_ctx *ctx, const zmpm_particles *views, int32_t nviews) {
    check_ctx(ctx);
    if (check_views(views, nviews) != ctx->n)
        throw ZmpmError(ZMPM_ERR_ARG, "export views must cover exactly the imported particles");
    ctx->set_device();
    int64_t off = 0;
    for (int k = 0; k < nviews; ++k) {
        launch_export(ctx->st[ctx->cur], ctx->n, &views[k], off, ctx->mats, ctx->aux(), ctx->stream);
        off += views[k].n;
    }
    HIP_CHECK(hipGetLastError());
    check_status(ctx);
}

}  // namespace

int32_t zmpm_check_particles(const zmpm_particles *view) {
    return guarded([&] { check_view(view); });
}

int32_t zmpm_import(zmpm_ctx *ctx, const zmpm_particles *view) {
    return guarded([&] { import_views(ctx, view, 1); });
}

int32_t zmpm_export(zmpm_ctx *ctx, const zmpm_particles *view) {
    return guarded([&] { export_views(ctx, view, 1); });
}

int32_t zmpm_import_objects(zmpm_ctx *ctx, const zmpm_particles *views, int32_t nviews) {
    return guarded([&] { import_views(ctx, views, nviews); });
}

int32_t zmpm_export_objects(zmpm_ctx *ctx, const zmpm_particles *views, int32_t nviews) {
    return guarded([&] { export_views(ctx, views, nviews); });
}

int64_t zmpm_num_particles(const zmpm_ctx *ctx) { return ctx ? ctx->n : 0; }

int32_t zmpm_particle_volume(zmpm_ctx *ctx, int32_t layout, int32_t tile_channels, float *vol, int32_t vol_chn,
                             float *mass, int32_t mass_chn) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->tiles_valid = false;  // the fused pass's tiles no longer match the state
        if (layout != ZMPM_LAYOUT_SOA && layout != ZMPM_LAYOUT_AOS && layout != ZMPM_LAYOUT_AOSOA32)
            throw ZmpmError(ZMPM_ERR_ARG, "unknown layout");
        if (layout == ZMPM_LAYOUT_AOSOA32 &&
            ((vol && (vol_chn < 0 || vol_chn >= tile_channels)) || (mass && (mass_chn < 0 || mass_chn >= tile_channels))))
            throw ZmpmError(ZMPM_ERR_ARG, "channel outside the tile");
        if (ctx->n == 0) return;
        ctx->set_device();
        if (!ctx->partitioned) full_partition(ctx);
        // the counts live in the grid's memory: the grid no longer holds a transfer
        launch_particle_volume(ctx->st[ctx->cur], ctx->n, ctx->part[ctx->pcur], ctx->L, ctx->mats,
                               reinterpret_cast<uint32_t *>(ctx->grid), ctx->max_gblocks(),
                               VolumeOut{layout, tile_channels, vol, vol_chn, mass, mass_chn}, ctx->status,
                               ctx->stream);
        HIP_CHECK(hipGetLastError());
        // the particles now carry these m, vol, which P2G reads (Maintenance.cu:50-51 -> Transfer.cu:857-858)
        ctx->pp_mass = true;
        ctx->sync_pp_flag();
        ctx->grid_ready = false;
        check_status(ctx);
    });
}

int32_t zmpm_partition(zmpm_ctx *ctx) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->tiles_valid = false;  // the fused pass's tiles no longer match the state
        ctx->set_device();
        full_partition(ctx);
    });
}

int32_t zmpm_set_transfer(zmpm_ctx *ctx, int32_t scheme) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->tiles_valid = false;  // the fused pass's tiles no longer match the state
        if (scheme != ZMPM_TRANSFER_APIC && scheme != ZMPM_TRANSFER_AFLIP && scheme != ZMPM_TRANSFER_FLIP)
            throw ZmpmError(ZMPM_ERR_ARG, "unknown transfer scheme");
        ctx->set_device();
        if (scheme != ZMPM_TRANSFER_APIC && !ctx->gstar) {
            ctx->tiles2 = (float *)ctx->alloc(sizeof(float) * 4 * kTileNodes * ctx->tiles_blocks);
            ctx->gstar = (float *)ctx->alloc(sizeof(float) * 256 * ctx->grid_blocks);
            HIP_CHECK(hipDeviceSynchronize());
        }
        // beta was not carried under APIC: the resident particles restart from their materials'
        if (scheme != ZMPM_TRANSFER_APIC && !(ctx->aux() & kAuxBeta) && ctx->n > 0) {
            launch_init_aux(ctx->st[ctx->cur], ctx->n, ctx->mats, kAuxBeta, ctx->stream);
            HIP_CHECK(hipGetLastError());
        }
        ctx->transfer = scheme;
        ctx->sync_pp_flag();
        ctx->grid_ready = false;
    });
}

int32_t zmpm_grid_clear(zmpm_ctx *ctx) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->tiles_valid = false;  // the fused pass's tiles no longer match the state
        if (!ctx->partitioned) throw ZmpmError(ZMPM_ERR_STATE, "zmpm_partition must precede zmpm_grid_clear");
        ctx->set_device();
        if (ctx->flip_style())
            launch_grid_aflip(GRID_CLEAR, ctx->part[ctx->pcur], ctx->L, ctx->tiles, ctx->tiles2, ctx->grid,
                              ctx->gstar, ctx->gvmax, ctx->max_gblocks(), ctx->stream);
        else
            launch_grid(GRID_CLEAR, ctx->part[ctx->pcur], ctx->L, ctx->tiles, ctx->grid, ctx->gvmax,
                        ctx->max_gblocks(), ctx->stream);
        HIP_CHECK(hipGetLastError());
        ctx->grid_ready = true;
        ctx->grid_part = ctx->pcur;
    });
}

int32_t zmpm_p2g(zmpm_ctx *ctx, float dt) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->tiles_valid = false;  // the fused pass's tiles no longer match the state
        if (!ctx->partitioned) throw ZmpmError(ZMPM_ERR_STATE, "zmpm_partition must precede zmpm_p2g");
        ctx->set_device();
        ctx->L.part_shift = ctx->psize[ctx->pcur];  // the parts the partition was cut into
        set_steps(ctx, step_args(dt, nullptr, 0, 0), SET_DT);
        launch_stress(ctx->st[ctx->cur], ctx->n, ctx->L, ctx->mats, ctx->sig, ctx->cap, ctx->stream,
                      ctx->stress_general());
        if (ctx->flip_style()) {
            if (ctx->transfer == ZMPM_TRANSFER_FLIP)
                launch_p2g_flip(ctx->st[ctx->cur], ctx->part[ctx->pcur], ctx->L, ctx->mats, ctx->sig, ctx->cap,
                                ctx->tiles, ctx->tiles2, ctx->status, ctx->max_units(), ctx->stream);
            else
                launch_p2g_aflip(ctx->st[ctx->cur], ctx->part[ctx->pcur], ctx->L, ctx->mats, ctx->sig, ctx->cap,
                                 ctx->tiles, ctx->tiles2, ctx->status, ctx->max_units(), ctx->stream);
            launch_grid_aflip(GRID_COMBINE, ctx->part[ctx->pcur], ctx->L, ctx->tiles, ctx->tiles2, ctx->grid,
                              ctx->gstar, ctx->gvmax, ctx->max_gblocks(), ctx->stream);
        } else {
            launch_p2g(ctx->st[ctx->cur], ctx->part[ctx->pcur], ctx->L, ctx->mats, ctx->sig, ctx->cap, ctx->tiles,
                       ctx->status, ctx->max_units(), ctx->stream, ctx->p2g_sparse, ctx->fused_path());
            launch_grid(GRID_COMBINE, ctx->part[ctx->pcur], ctx->L, ctx->tiles, ctx->grid, ctx->gvmax,
                        ctx->max_gblocks(), ctx->stream);
        }
        HIP_CHECK(hipGetLastError());
        ctx->grid_ready = true;
        ctx->grid_part = ctx->pcur;
    });
}

int32_t zmpm_grid_update(zmpm_ctx *ctx, float dt, const float accel[3]) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->tiles_valid = false;  // the fused pass's tiles no longer match the state
        if (!ctx->grid_ready) throw ZmpmError(ZMPM_ERR_STATE, "zmpm_p2g must precede zmpm_grid_update");
        ctx->set_device();
        const float g0[3] = {0.0f, -9.8f, 0.0f};  // UpdateZSGrid default gravity (MPMPipeline.cu:947)
        set_steps(ctx, step_args(dt, accel ? accel : g0, 0, 0), SET_DT | SET_ACCEL);
        if (ctx->flip_style())
            launch_grid_aflip(GRID_UPDATE, ctx->part[ctx->pcur], ctx->L, ctx->tiles, ctx->tiles2, ctx->grid,
                              ctx->gstar, ctx->gvmax, ctx->max_gblocks(), ctx->stream);
        else
            launch_grid(GRID_UPDATE, ctx->part[ctx->pcur], ctx->L, ctx->tiles, ctx->grid, ctx->gvmax,
                        ctx->max_gblocks(), ctx->stream);
        HIP_CHECK(hipGetLastError());
    });
}

int32_t zmpm_boundary(zmpm_ctx *ctx, int32_t boundary, int32_t pad) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->tiles_valid = false;  // the fused pass's tiles no longer match the state
        if (!ctx->grid_ready) throw ZmpmError(ZMPM_ERR_STATE, "zmpm_p2g must precede zmpm_boundary");
        if (boundary != ZMPM_BOUNDARY_NONE && boundary != ZMPM_BOUNDARY_STICKY_BOX)
            throw ZmpmError(ZMPM_ERR_ARG, "unknown boundary type");
        if (boundary == ZMPM_BOUNDARY_NONE && ctx->L.ncols == 0 && ctx->bnd.n == 0) return;
        ctx->set_device();
        set_steps(ctx, step_args(0.0f, nullptr, boundary, pad), SET_BOUNDARY);
        const bool aflip = ctx->flip_style();  // flip style: on vstar
        if (aflip)
            launch_grid_aflip(GRID_BOUNDARY, ctx->part[ctx->pcur], ctx->L, ctx->tiles, ctx->tiles2, ctx->grid,
                              ctx->gstar, ctx->gvmax, ctx->max_gblocks(), ctx->stream);
        else
            launch_grid(GRID_BOUNDARY, ctx->part[ctx->pcur], ctx->L, ctx->tiles, ctx->grid, ctx->gvmax,
                        ctx->max_gblocks(), ctx->stream);
        launch_colliders(ctx->part[ctx->pcur], ctx->L, aflip ? ctx->gstar : ctx->grid, ctx->gvmax, ctx->max_gblocks(),
                         false, ctx->stream);
        launch_boundary_particles(ctx->bnd, ctx->part[ctx->pcur], ctx->L, aflip ? ctx->gstar : ctx->grid, nullptr,
                                  ctx->max_gblocks(), ctx->stream);
        HIP_CHECK(hipGetLastError());
    });
}

int32_t zmpm_g2p(zmpm_ctx *ctx, float dt) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->tiles_valid = false;  // the fused pass's tiles no longer match the state
        if (!ctx->grid_ready) throw ZmpmError(ZMPM_ERR_STATE, "zmpm_p2g must precede zmpm_g2p");
        ctx->set_device();
        ctx->L.part_shift = ctx->psize[ctx->pcur];  // the parts the partition was cut into
        set_steps(ctx, step_args(dt, nullptr, 0, 0), SET_DT);
        State A = ctx->st[ctx->cur];
        const bool aflip = ctx->flip_style();
        launch_g2p_update(false, A, A, ctx->part[ctx->pcur], ctx->part[ctx->pcur], ctx->L, ctx->grid, ctx->bin,
                          ctx->max_units(), ctx->stream, 256, aflip ? ctx->gstar : nullptr, ctx->aux(),
                          ctx->transfer == ZMPM_TRANSFER_FLIP);
        HIP_CHECK(hipGetLastError());
        ctx->partitioned = false;  // particles moved: the next P2G needs a fresh partition
        ctx->grid_ready = false;
    });
}

namespace {
// zmpm_collider -> DevCollider: validated, half-space normal made unit, R from yaw/pitch/roll.
DevCollider dev_collider(const zmpm_collider &c, int32_t nscenes) {
    if (c.type < ZMPM_COLLIDER_STICKY || c.type > ZMPM_COLLIDER_SEPARATE)
        throw ZmpmError(ZMPM_ERR_ARG, "unknown collider type");
    if (c.shape < ZMPM_SHAPE_HALFSPACE || c.shape > ZMPM_SHAPE_SDF_LEAVES)
        throw ZmpmError(ZMPM_ERR_ARG, "unknown collider shape");
    if (c.scene < 0 || c.scene >= nscenes) throw ZmpmError(ZMPM_ERR_ARG, "collider scene out of range");
    if (!(c.scale > 0.0f) || !std::isfinite(c.scale)) throw ZmpmError(ZMPM_ERR_ARG, "collider scale must be > 0");
    DevCollider d{};
    d.type = c.type;
    d.shape = c.shape;
    d.scene = c.scene;
    for (int k = 0; k < 6; ++k) d.shape_params[k] = c.shape_params[k];
    if (c.shape == ZMPM_SHAPE_HALFSPACE) {
        const double nn = std::sqrt((double)c.shape_params[0] * c.shape_params[0] +
                                    (double)c.shape_params[1] * c.shape_params[1] +
                                    (double)c.shape_params[2] * c.shape_params[2]);
        if (!(nn > 0.0)) throw ZmpmError(ZMPM_ERR_ARG, "half-space normal is zero");
        for (int k = 0; k < 4; ++k) d.shape_params[k] = (float)(c.shape_params[k] / nn);
    } else if (c.shape == ZMPM_SHAPE_SPHERE) {
        if (!(c.shape_params[3] >= 0.0f)) throw ZmpmError(ZMPM_ERR_ARG, "sphere radius must be >= 0");
    } else if (c.shape == ZMPM_SHAPE_BOX) {
        for (int k = 0; k < 3; ++k)
            if (!(c.shape_params[k] <= c.shape_params[3 + k])) throw ZmpmError(ZMPM_ERR_ARG, "box lo > hi");
    } else if (c.shape == ZMPM_SHAPE_SDF_GRID) {
        if (!c.sdf) throw ZmpmError(ZMPM_ERR_ARG, "SDF grid without samples");
        for (int k = 0; k < 3; ++k)
            if (c.dims[k] < 2) throw ZmpmError(ZMPM_ERR_ARG, "SDF grid needs >= 2 samples per axis");
        if (!(c.spacing > 0.0f)) throw ZmpmError(ZMPM_ERR_ARG, "SDF spacing must be > 0");
    } else {
        if (!c.sdf || !c.leaf_origins) throw ZmpmError(ZMPM_ERR_ARG, "SDF leaves without samples or origins");
        if (c.nleaves < 1 || c.nleaves > kMaxSdfLeaves)
            throw ZmpmError(ZMPM_ERR_ARG, "SDF leaf count must be in [1, 2^24]");
        if (!(c.spacing > 0.0f) || !std::isfinite(c.spacing)) throw ZmpmError(ZMPM_ERR_ARG, "SDF spacing must be > 0");
        if (!std::isfinite(c.shape_params[0])) throw ZmpmError(ZMPM_ERR_ARG, "SDF background must be finite");
    }
    for (int k = 0; k < 3; ++k) {
        d.dims[k] = c.dims[k];
        d.origin[k] = c.origin[k];
        d.b[k] = c.translation[k];
        d.dbdt[k] = c.translation_rate[k];
        d.omega[k] = c.angular_velocity[k];
    }
    d.spacing = c.spacing;
    d.sdf = c.sdf;
    d.s = c.scale;
    d.dsdt = c.scale_rate;
    // R = Ry(yaw) Rx(pitch) Rz(roll), row-major, collider frame -> world
    const double r = 3.14159265358979323846 / 180.0;
    const double cy = std::cos(c.ypr_degrees[0] * r), sy = std::sin(c.ypr_degrees[0] * r);
    const double cp = std::cos(c.ypr_degrees[1] * r), sp = std::sin(c.ypr_degrees[1] * r);
    const double cr = std::cos(c.ypr_degrees[2] * r), sr = std::sin(c.ypr_degrees[2] * r);
    const double Ry[3][3] = {{cy, 0, sy}, {0, 1, 0}, {-sy, 0, cy}};
    const double Rx[3][3] = {{1, 0, 0}, {0, cp, -sp}, {0, sp, cp}};
    const double Rz[3][3] = {{cr, -sr, 0}, {sr, cr, 0}, {0, 0, 1}};
    double T[3][3], R[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) T[i][j] = Ry[i][0] * Rx[0][j] + Ry[i][1] * Rx[1][j] + Ry[i][2] * Rx[2][j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) R[i][j] = T[i][0] * Rz[0][j] + T[i][1] * Rz[1][j] + T[i][2] * Rz[2][j];
    for (int i = 0; i < 9; ++i) d.R[i] = (float)R[i / 3][i % 3];
    return d;
}

// The dense leaf map of an SDF_LEAVES collider: the caller's leaf origins read once, checked
// (multiples of 8, distinct), and every 8^3 box of their bounding box mapped to its leaf or -1.
// The map stays allocated until the context is destroyed: a graph captured earlier may still
// read a collider table entry that points at it.
void build_leaf_map(zmpm_ctx *ctx, const zmpm_collider &c, DevCollider &d) {
    std::vector<int32_t> org((size_t)c.nleaves * 3);
    HIP_CHECK(hipMemcpyAsync(org.data(), c.leaf_origins, org.size() * sizeof(int32_t), hipMemcpyDefault,
                             ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    int64_t lo[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, hi[3] = {INT64_MIN, INT64_MIN, INT64_MIN};
    for (int32_t l = 0; l < c.nleaves; ++l)
        for (int k = 0; k < 3; ++k) {
            const int32_t o = org[3 * (size_t)l + k];
            if (o % 8 != 0) throw ZmpmError(ZMPM_ERR_ARG, "SDF leaf origins must be multiples of 8");
            lo[k] = std::min<int64_t>(lo[k], o >> 3);
            hi[k] = std::max<int64_t>(hi[k], o >> 3);
        }
    int64_t dims[3], total = 1;
    for (int k = 0; k < 3; ++k) {
        dims[k] = hi[k] - lo[k] + 1;
        total *= dims[k];
        if (total > kMaxSdfLeafMap)
            throw ZmpmError(ZMPM_ERR_ARG, "SDF leaves' bounding box exceeds 2^28 leaf boxes");
    }
    std::vector<int32_t> map((size_t)total, -1);
    for (int32_t l = 0; l < c.nleaves; ++l) {
        const int32_t *o = &org[3 * (size_t)l];
        const int64_t slot = (((o[0] >> 3) - lo[0]) * dims[1] + ((o[1] >> 3) - lo[1])) * dims[2] + ((o[2] >> 3) - lo[2]);
        if (map[(size_t)slot] >= 0) throw ZmpmError(ZMPM_ERR_ARG, "SDF leaf origins must be distinct");
        map[(size_t)slot] = l;
    }
    int32_t *dmap = (int32_t *)ctx->alloc(map.size() * sizeof(int32_t));
    HIP_CHECK(hipMemcpyAsync(dmap, map.data(), map.size() * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    d.leaf_map = dmap;
    for (int k = 0; k < 3; ++k) {
        d.leaf_lo[k] = (int32_t)lo[k];
        d.dims[k] = (int32_t)dims[k];
    }
}
}  // namespace

int32_t zmpm_set_colliders(zmpm_ctx *ctx, const zmpm_collider *cols, int32_t count) {
    return guarded([&] {
        check_ctx(ctx);
        if (count < 0 || count > kMaxColliders) throw ZmpmError(ZMPM_ERR_ARG, "collider count must be in [0, 64]");
        if (count > 0 && !cols) throw ZmpmError(ZMPM_ERR_ARG, "null collider array");
        std::vector<DevCollider> dc;
        bool moving = false;
        for (int i = 0; i < count; ++i) dc.push_back(dev_collider(cols[i], ctx->L.nscenes));
        ctx->set_device();
        for (int i = 0; i < count; ++i) {
            const zmpm_collider &c = cols[i];
            if (c.shape == ZMPM_SHAPE_SDF_LEAVES) build_leaf_map(ctx, c, dc[(size_t)i]);
            moving = moving || c.translation_rate[0] != 0.0f || c.translation_rate[1] != 0.0f ||
                     c.translation_rate[2] != 0.0f || c.scale_rate != 0.0f;
        }
        if (!ctx->L.cols) ctx->L.cols = (DevCollider *)ctx->alloc(sizeof(DevCollider) * kMaxColliders);
        if (count > 0)
            HIP_CHECK(hipMemcpyAsync(ctx->L.cols, dc.data(), sizeof(DevCollider) * count, hipMemcpyHostToDevice,
                                     ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        ctx->L.ncols = count;
        ctx->cols_move = moving;
    });
}

int32_t zmpm_set_boundary_particles(zmpm_ctx *ctx, const zmpm_boundary_particles *bp, int32_t type) {
    return guarded([&] {
        check_ctx(ctx);
        if (type < ZMPM_COLLIDER_STICKY || type > ZMPM_COLLIDER_SEPARATE)
            throw ZmpmError(ZMPM_ERR_ARG, "unknown boundary response type");
        if (!bp || bp->n == 0) {
            ctx->bnd.n = 0;
            return;
        }
        if (bp->n < 0 || !bp->x || !bp->v || !bp->m) throw ZmpmError(ZMPM_ERR_ARG, "invalid boundary particles");
        if (!bp->nrm && type != ZMPM_COLLIDER_STICKY)
            throw ZmpmError(ZMPM_ERR_ARG, "slip and separate boundaries need normals");
        if (bp->scene < 0 || bp->scene >= ctx->L.nscenes) throw ZmpmError(ZMPM_ERR_ARG, "boundary scene out of range");
        ctx->set_device();
        if (!ctx->bnd.acc) {
            // 7 int64 sums per node of every grid block the partition can hold, and the scale word
            const int64_t blocks = std::min<int64_t>(ctx->L.size, 8 * std::max<int64_t>(ctx->cap, 1));
            ctx->bnd.acc = (unsigned long long *)ctx->alloc(sizeof(unsigned long long) * kBndChannels * 64 * blocks);
            ctx->bnd.scale_bits = (uint32_t *)ctx->alloc(sizeof(uint32_t));
            HIP_CHECK(hipDeviceSynchronize());
        }
        ctx->bnd.n = bp->n;
        ctx->bnd.x = bp->x;
        ctx->bnd.v = bp->v;
        ctx->bnd.m = bp->m;
        ctx->bnd.nrm = bp->nrm;
        ctx->bnd.scene = bp->scene;
        ctx->bnd.type = type;
    });
}

int32_t zmpm_step_colliders(zmpm_ctx *ctx, float dt) {
    return guarded([&] {
        check_ctx(ctx);
        if (!(dt >= 0.0f)) throw ZmpmError(ZMPM_ERR_ARG, "dt must be >= 0");
        ctx->set_device();
        launch_step_colliders(ctx->L, dt, ctx->stream);
        HIP_CHECK(hipGetLastError());
    });
}

int32_t zmpm_return_mapping(zmpm_ctx *ctx) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->tiles_valid = false;  // the fused pass's tiles no longer match the state
        ctx->set_device();
        if (ctx->has_plastic) launch_return_mapping(ctx->st[ctx->cur], ctx->n, ctx->mats, ctx->stream);
        HIP_CHECK(hipGetLastError());
    });
}

int32_t zmpm_compute_particle_beta(zmpm_ctx *ctx, const zmpm_step_params *p, float jp_critical) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->tiles_valid = false;  // the fused pass's tiles no longer match the state
        if (p && !(p->dt >= 0.0f)) throw ZmpmError(ZMPM_ERR_ARG, "dt must be >= 0");
        if (!std::isfinite(jp_critical)) throw ZmpmError(ZMPM_ERR_ARG, "jp_critical must be finite");
        ctx->set_device();
        if (ctx->bnd.n > 0) {
            // the boundary grid (Boundary.cu:173-219): the boundary particles' sums on the current partition
            if (!ctx->partitioned) full_partition(ctx);
            launch_boundary_particles(ctx->bnd, ctx->part[ctx->pcur], ctx->L, nullptr, nullptr, ctx->max_gblocks(),
                                      ctx->stream, false);
        }
        // the scene table's dt (its entry 0 travels by value) is the one the kernel reads when p is NULL
        launch_particle_beta(ctx->st[ctx->cur], ctx->n, ctx->L, ctx->mats, p ? p->dt : -1.0f, jp_critical, ctx->bnd,
                             ctx->part[ctx->pcur], ctx->stream);
        HIP_CHECK(hipGetLastError());
        ctx->beta_kept = true;
    });
}

int32_t zmpm_scene_max_vel_sqr(zmpm_ctx *ctx, float *out, int32_t nscenes) {
    return guarded([&] {
        check_ctx(ctx);
        if (!out || nscenes < 0) throw ZmpmError(ZMPM_ERR_ARG, "invalid output");
        ctx->set_device();
        const int k = std::min<int>(nscenes, ctx->L.nscenes);
        reduce_vmax(ctx);
        std::vector<uint32_t> bits(ctx->L.nscenes);
        HIP_CHECK(hipMemcpyAsync(bits.data(), ctx->vmax, sizeof(uint32_t) * bits.size(), hipMemcpyDeviceToHost,
                                 ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < k; ++i) std::memcpy(&out[i], &bits[i], sizeof(float));
        check_status(ctx);
    });
}

int32_t zmpm_max_vel_sqr_device(zmpm_ctx *ctx, const float **out) {
    return guarded([&] {
        check_ctx(ctx);
        if (!out) throw ZmpmError(ZMPM_ERR_ARG, "null output");
        *out = reinterpret_cast<const float *>(ctx->vmax);
    });
}

int32_t zmpm_max_vel_sqr(zmpm_ctx *ctx, float *out) {
    return guarded([&] {
        check_ctx(ctx);
        if (!out) throw ZmpmError(ZMPM_ERR_ARG, "null output");
        ctx->set_device();
        reduce_vmax(ctx);
        std::vector<uint32_t> bits(ctx->L.nscenes);
        HIP_CHECK(hipMemcpyAsync(bits.data(), ctx->vmax, sizeof(uint32_t) * bits.size(), hipMemcpyDeviceToHost,
                                 ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        uint32_t mx = 0;  // non-negative floats order as their bits
        for (uint32_t b : bits) mx = std::max(mx, b);
        std::memcpy(out, &mx, sizeof(float));
        check_status(ctx);
    });
}

namespace {

constexpr int kPhases = 6;  // stress, P2G, grid, G2P move, partition, G2P update

// One fused substep on the resident, partitioned state, every scene with the
// step parameters in its scene-table entry.  ev (optional) receives kPhases + 1
// events recorded around the phases on the context stream.  grow > 0 sizes the
// launches for a graph replayed later (block counts may change by then).
// The fused chain (APIC): the tiles of the current state come from the previous
// substep's fused pass (or, first, from the stress + P2G kernels), then the grid
// stages, the move pass, the lattice scan and the fused G2P / P2G pass, which
// leaves the tiles of the next state.  last: the call's final substep, which
// also stores v and C (no kernel of the chain reads them in between).
void substep_fused(zmpm_ctx *ctx, hipEvent_t *ev, int grow, bool last) {
    hipStream_t s = ctx->stream;
    const State A = ctx->st[ctx->cur], B = ctx->st[ctx->cur ^ 1];
    const Partition P = ctx->part[ctx->pcur], Pn = ctx->part[ctx->pcur ^ 1];
    const int grid_mode = GRID_COMBINE | GRID_UPDATE | GRID_BOUNDARY;
    const int64_t npb = ctx->max_pblocks() << grow, ngb = ctx->max_gblocks() << grow;
    if (ev) HIP_CHECK(hipEventRecord(ev[0], s));
    if (!ctx->tiles_valid) launch_stress(A, ctx->n, ctx->L, ctx->mats, ctx->sig, ctx->cap, s, ctx->stress_general());
    if (ev) HIP_CHECK(hipEventRecord(ev[1], s));
    if (!ctx->tiles_valid)
        launch_p2g(A, P, ctx->L, ctx->mats, ctx->sig, ctx->cap, ctx->tiles, ctx->status, npb, s, ctx->p2g_sparse, true);
    if (ev) HIP_CHECK(hipEventRecord(ev[2], s));
    launch_grid(grid_mode, P, ctx->L, ctx->tiles, ctx->grid, ctx->gvmax, ngb, s, nullptr, ctx->vmax, ctx->L.nscenes);
    launch_colliders(P, ctx->L, ctx->grid, ctx->gvmax, ngb, true, s);
    launch_boundary_particles(ctx->bnd, P, ctx->L, ctx->grid, ctx->gvmax, ngb, s);
    if (ctx->cols_move) launch_step_colliders(ctx->L, -1.0f, s);
    if (ev) HIP_CHECK(hipEventRecord(ev[3], s));
    // the move pass also reduces this substep's MaxVelSqr (ctx->bin.vmax_out)
    launch_g2p_move(A, P, ctx->L, ctx->grid, ctx->gvmax, ctx->bin, ctx->status, npb, s, ctx->wg_move, nullptr, false,
                    true);
    if (ev) HIP_CHECK(hipEventRecord(ev[4], s));
    launch_lattice_scan(ctx->L, Pn, ctx->bin, P.tile_used, ctx->n, ctx->status, s);
    ctx->psize[ctx->pcur ^ 1] = ctx->L.part_shift;
    if (ev) HIP_CHECK(hipEventRecord(ev[5], s));
    // the new partition's blocks: at most the old count plus what the lattice scan adds, as G2P update sizes it
    launch_g2p2g(A, B, P, Pn, ctx->L, ctx->grid, ctx->bin, ctx->arr, ctx->tiles, ctx->mats, ctx->status, ctx->aux(),
                 last, ctx->has_plastic || ctx->has_stvk || ctx->has_aniso, ctx->n, npb, s);
    if (ev) HIP_CHECK(hipEventRecord(ev[6], s));
    ctx->tiles_valid = true;
    ctx->grid_part = ctx->pcur;
    ctx->cur ^= 1;
    ctx->pcur ^= 1;
}

// Kernels read a partition's parts with the part shift it was cut with, which every launch
// passes in L.part_shift: the current partition, cut into parts of 2^shift particles (the
// fused chain: kNoSplit, one part per block).  A partition cut otherwise (by the other chain,
// before the fused pass was switched) is rebuilt from the current, block-sorted state.
void align_parts(zmpm_ctx *ctx, uint32_t shift) {
    if (!ctx->partitioned) {
        full_partition(ctx, shift);
    } else if (ctx->psize[ctx->pcur] != shift) {
        ctx->loaded_sorted = true;
        full_partition(ctx, shift);
        ctx->tiles_valid = false;
    }
    ctx->L.part_shift = shift;
}

// The tiles of the current state for the fused chain (stress + P2G), if it runs and they are
// stale.  force: a fused graph is about to run whatever the context's current setting.
void fused_prologue(zmpm_ctx *ctx, bool force = false) {
    if (!(ctx->fused_path() || force) || ctx->tiles_valid) return;
    const State A = ctx->st[ctx->cur];
    const Partition P = ctx->part[ctx->pcur];
    launch_stress(A, ctx->n, ctx->L, ctx->mats, ctx->sig, ctx->cap, ctx->stream, ctx->stress_general());
    launch_p2g(A, P, ctx->L, ctx->mats, ctx->sig, ctx->cap, ctx->tiles, ctx->status, ctx->max_units(), ctx->stream,
               ctx->p2g_sparse, true);
    ctx->tiles_valid = true;
}

void substep_once(zmpm_ctx *ctx, hipEvent_t *ev, int grow = 0, bool last = true) {
    if (ctx->fused_path()) {
        substep_fused(ctx, ev, grow, last);
        return;
    }
    hipStream_t s = ctx->stream;
    const State A = ctx->st[ctx->cur], B = ctx->st[ctx->cur ^ 1];
    const Partition P = ctx->part[ctx->pcur], Pn = ctx->part[ctx->pcur ^ 1];
    const int grid_mode = GRID_COMBINE | GRID_UPDATE | GRID_BOUNDARY;
    // P2G and both G2P passes work per part
    const int64_t npb = ctx->max_units() << grow, ngb = ctx->max_gblocks() << grow;
    ctx->tiles_valid = false;
    if (ev) HIP_CHECK(hipEventRecord(ev[0], s));
    launch_stress(A, ctx->n, ctx->L, ctx->mats, ctx->sig, ctx->cap, s, ctx->stress_general());
    if (ev) HIP_CHECK(hipEventRecord(ev[1], s));
    // FLIP transfers: momentum and stress apart, velocities v and vstar
    const bool aflip = ctx->flip_style(), nonaffine = ctx->transfer == ZMPM_TRANSFER_FLIP;
    float *gstar = aflip ? ctx->gstar : nullptr;
    if (nonaffine)
        launch_p2g_flip(A, P, ctx->L, ctx->mats, ctx->sig, ctx->cap, ctx->tiles, ctx->tiles2, ctx->status, npb, s);
    else if (aflip)
        launch_p2g_aflip(A, P, ctx->L, ctx->mats, ctx->sig, ctx->cap, ctx->tiles, ctx->tiles2, ctx->status, npb, s);
    else
        launch_p2g(A, P, ctx->L, ctx->mats, ctx->sig, ctx->cap, ctx->tiles, ctx->status, npb, s, ctx->p2g_sparse);
    if (ev) HIP_CHECK(hipEventRecord(ev[2], s));
    if (aflip)
        launch_grid_aflip(grid_mode, P, ctx->L, ctx->tiles, ctx->tiles2, ctx->grid, gstar, ctx->gvmax, ngb, s,
                          ctx->vmax, ctx->L.nscenes);
    else
        launch_grid(grid_mode, P, ctx->L, ctx->tiles, ctx->grid, ctx->gvmax, ngb, s, nullptr, ctx->vmax,
                    ctx->L.nscenes);
    launch_colliders(P, ctx->L, aflip ? gstar : ctx->grid, ctx->gvmax, ngb, true, s);  // level-set colliders, if any
    launch_boundary_particles(ctx->bnd, P, ctx->L, aflip ? gstar : ctx->grid, ctx->gvmax, ngb, s);  // if any
    if (ctx->cols_move) launch_step_colliders(ctx->L, -1.0f, s);  // StepZSBoundary by each scene's dt
    if (ev) HIP_CHECK(hipEventRecord(ev[3], s));
    // G2P pass 1: direction of every particle's new block, per-block counts, destination marks, and
    // UpdateZSGrid's MaxVelSqr output of this substep (MPMPipeline.cu:939-942) per scene on the
    // device, into the words k_grid cleared
    launch_g2p_move(A, P, ctx->L, ctx->grid, ctx->gvmax, ctx->bin, ctx->status, npb, s, ctx->wg_move, gstar,
                    nonaffine);
    if (ev) HIP_CHECK(hipEventRecord(ev[4], s));
    // the next partition: particle blocks, their first particles, grid blocks
    launch_lattice_scan(ctx->L, Pn, ctx->bin, P.tile_used, ctx->n, ctx->status, s);
    ctx->psize[ctx->pcur ^ 1] = ctx->L.part_shift;
    if (ev) HIP_CHECK(hipEventRecord(ev[5], s));
    // G2P pass 2: full update, written straight into the next stable block order
    launch_g2p_update(true, A, B, P, Pn, ctx->L, ctx->grid, ctx->bin, npb, s, ctx->wg_update, gstar, ctx->aux(),
                      nonaffine);
    if (ctx->has_plastic) launch_return_mapping(B, ctx->n, ctx->mats, s);  // ZSReturnMapping after G2P
    if (ev) HIP_CHECK(hipEventRecord(ev[6], s));
    ctx->grid_part = ctx->pcur;
    ctx->cur ^= 1;
    ctx->pcur ^= 1;
}

void check_step_params(const zmpm_step_params *p, int32_t nsteps) {
    if (nsteps < 0) throw ZmpmError(ZMPM_ERR_ARG, "nsteps must be >= 0");
    if (p && p->boundary != ZMPM_BOUNDARY_NONE && p->boundary != ZMPM_BOUNDARY_STICKY_BOX)
        throw ZmpmError(ZMPM_ERR_ARG, "unknown boundary type");
}

// params (optional) -> every scene's step parameters
void apply_step_params(zmpm_ctx *ctx, const zmpm_step_params *p) {
    if (p) set_steps(ctx, step_args(p->dt, p->gravity, p->boundary, p->pad), SET_DT | SET_ACCEL | SET_BOUNDARY);
}

}  // namespace

int32_t zmpm_substeps(zmpm_ctx *ctx, const zmpm_step_params *p, int32_t nsteps) {
    return guarded([&] {
        check_ctx(ctx);
        check_step_params(p, nsteps);
        ctx->set_device();
        align_parts(ctx, ctx->wanted_part_shift());
        apply_step_params(ctx, p);
        for (int step = 0; step < nsteps; ++step) substep_once(ctx, nullptr, 0, step == nsteps - 1);
        HIP_CHECK(hipGetLastError());
        ctx->partitioned = true;
        ctx->grid_ready = false;
    });
}

int32_t zmpm_substeps_timed(zmpm_ctx *ctx, const zmpm_step_params *p, int32_t nsteps, float *phase_ms,
                            int32_t nphases) {
    return guarded([&] {
        check_ctx(ctx);
        check_step_params(p, nsteps);
        if (!phase_ms || nphases < 1) throw ZmpmError(ZMPM_ERR_ARG, "phase_ms must hold at least one entry");
        ctx->set_device();
        align_parts(ctx, ctx->wanted_part_shift());
        apply_step_params(ctx, p);
        std::vector<hipEvent_t> ev((size_t)nsteps * (kPhases + 1));
        for (auto &e : ev) HIP_CHECK(hipEventCreate(&e));
        for (int step = 0; step < nsteps; ++step)
            substep_once(ctx, &ev[(size_t)step * (kPhases + 1)], 0, step == nsteps - 1);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        double acc[kPhases] = {0};
        for (int step = 0; step < nsteps; ++step)
            for (int k = 0; k < kPhases; ++k) {
                float ms = 0.0f;
                HIP_CHECK(hipEventElapsedTime(&ms, ev[(size_t)step * (kPhases + 1) + k],
                                              ev[(size_t)step * (kPhases + 1) + k + 1]));
                acc[k] += ms;
            }
        for (auto &e : ev) (void)hipEventDestroy(e);
        for (int k = 0; k < nphases; ++k) phase_ms[k] = k < kPhases ? (float)acc[k] : 0.0f;
        ctx->partitioned = true;
        ctx->grid_ready = false;
    });
}

int32_t zmpm_autotune(zmpm_ctx *ctx, const zmpm_step_params *p, int32_t nsteps, int32_t *choice) {
    return guarded([&] {
        check_ctx(ctx);
        check_step_params(p, nsteps);
        if (nsteps <= 0) throw ZmpmError(ZMPM_ERR_ARG, "nsteps must be positive");
        ctx->set_device();
        if (!ctx->partitioned) full_partition(ctx);
        // every candidate runs from the same saved state; results do not depend on the choice
        void *saved = nullptr;
        HIP_CHECK(hipMalloc(&saved, sizeof(float) * std::max<int64_t>(state_floats(ctx->n), 64)));
        hipEvent_t e0 = nullptr, e1 = nullptr;
        try {
            HIP_CHECK(hipMemcpyAsync(saved, ctx->st[ctx->cur].ch, sizeof(float) * state_floats(ctx->n),
                                     hipMemcpyDeviceToDevice, ctx->stream));
            HIP_CHECK(hipEventCreate(&e0));
            HIP_CHECK(hipEventCreate(&e1));
            // moving colliders advance in every substep: each run starts from their saved frames
            std::vector<DevCollider> cols(ctx->L.ncols);
            if (ctx->L.ncols > 0) {
                HIP_CHECK(hipMemcpyAsync(cols.data(), ctx->L.cols, sizeof(DevCollider) * ctx->L.ncols,
                                         hipMemcpyDeviceToHost, ctx->stream));
                HIP_CHECK(hipStreamSynchronize(ctx->stream));
            }
            auto restore_colliders = [&] {
                if (ctx->L.ncols > 0)
                    HIP_CHECK(hipMemcpyAsync(ctx->L.cols, cols.data(), sizeof(DevCollider) * ctx->L.ncols,
                                             hipMemcpyHostToDevice, ctx->stream));
            };
            apply_step_params(ctx, p);
            // G2P sizes with the sparse P2G on, then the sparse P2G off with the best sizes, then (APIC)
            // the fused G2P / P2G pass with either move size
            struct Cand {
                int move, update, sparse, fuse;
            };
            Cand cand[7] = {{256, 256, 1, 0}, {64, 256, 1, 0}, {256, 64, 1, 0}, {64, 64, 1, 0},
                            {0, 0, 0, 0},     {256, 0, 1, 1},  {64, 0, 1, 1}};
            const int ncand = ctx->transfer == ZMPM_TRANSFER_APIC ? 7 : 5;
            float best = -1.0f;
            int pick = 0;
            for (int k = 0; k < ncand; ++k) {
                if (k == 4) cand[4] = Cand{cand[pick].move, cand[pick].update, 0, 0};
                ctx->wg_move = cand[k].move;
                if (!cand[k].fuse) ctx->wg_update = cand[k].update;
                ctx->p2g_sparse = cand[k].sparse != 0;
                ctx->fuse = cand[k].fuse != 0;
                HIP_CHECK(hipMemcpyAsync(ctx->st[ctx->cur].ch, saved, sizeof(float) * state_floats(ctx->n),
                                         hipMemcpyDeviceToDevice, ctx->stream));
                restore_colliders();
                ctx->loaded_sorted = true;
                full_partition(ctx);
                substep_once(ctx, nullptr);  // warm-up (also the fused chain's first P2G)
                HIP_CHECK(hipEventRecord(e0, ctx->stream));
                for (int step = 0; step < nsteps; ++step) substep_once(ctx, nullptr, 0, step == nsteps - 1);
                HIP_CHECK(hipEventRecord(e1, ctx->stream));
                HIP_CHECK(hipEventSynchronize(e1));
                float ms = 0.0f;
                HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                if (best < 0.0f || ms < best) {
                    best = ms;
                    pick = k;
                }
            }
            ctx->wg_move = cand[pick].move;
            if (!cand[pick].fuse) ctx->wg_update = cand[pick].update;
            ctx->p2g_sparse = cand[pick].sparse != 0;
            ctx->fuse = cand[pick].fuse != 0;
            HIP_CHECK(hipMemcpyAsync(ctx->st[ctx->cur].ch, saved, sizeof(float) * state_floats(ctx->n),
                                     hipMemcpyDeviceToDevice, ctx->stream));
            restore_colliders();
            ctx->loaded_sorted = true;
            full_partition(ctx);
            HIP_CHECK(hipStreamSynchronize(ctx->stream));
            if (choice) {
                choice[0] = ctx->wg_move;
                choice[1] = ctx->fused_path() ? 0 : ctx->wg_update;
                choice[2] = ctx->p2g_sparse ? 1 : 0;
            }
        } catch (...) {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipFree(saved);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        HIP_CHECK(hipFree(saved));
        check_status(ctx);
    });
}

int32_t zmpm_graph_capture(zmpm_ctx *ctx, const zmpm_step_params *p, int32_t nsteps) {
    return guarded([&] {
        check_ctx(ctx);
        check_step_params(p, nsteps);
        if (nsteps <= 0 || (nsteps & 1))
            throw ZmpmError(ZMPM_ERR_ARG, "a captured chain needs an even number of substeps (double buffers)");
        ctx->set_device();
        align_parts(ctx, ctx->wanted_part_shift());
        ctx->graph_shift = ctx->L.part_shift;
        // the fused chain's first tiles, eagerly, so that the graph holds no stress + P2G prologue
        apply_step_params(ctx, p);
        fused_prologue(ctx);
        const bool tiles_valid = ctx->tiles_valid;
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (ctx->graph_exec) {
            (void)hipGraphExecDestroy(ctx->graph_exec);
            ctx->graph_exec = nullptr;
        }
        if (ctx->graph) {
            (void)hipGraphDestroy(ctx->graph);
            ctx->graph = nullptr;
        }
        if (!ctx->capture_stream) HIP_CHECK(hipStreamCreateWithFlags(&ctx->capture_stream, hipStreamNonBlocking));
        // capture on a stream of our own (the context stream may be the legacy default stream)
        const hipStream_t user = ctx->stream;
        const int cur = ctx->cur, pcur = ctx->pcur;
        ctx->stream = ctx->capture_stream;
        hipError_t err = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
        if (err == hipSuccess) {
            try {
                apply_step_params(ctx, p);
                // launches sized for up to 4x the current block counts; kernels loop over any excess
                for (int step = 0; step < nsteps; ++step) substep_once(ctx, nullptr, 2, step == nsteps - 1);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(ctx->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                ctx->stream = user;
                ctx->cur = cur;
                ctx->pcur = pcur;
                throw;
            }
            err = hipStreamEndCapture(ctx->stream, &ctx->graph);
        }
        ctx->stream = user;
        ctx->cur = cur;  // an even number of substeps leaves the buffers where they were
        ctx->pcur = pcur;
        ctx->tiles_valid = tiles_valid;  // nothing ran during the capture
        ctx->graph_fused = ctx->fused_path();
        HIP_CHECK(err);
        HIP_CHECK(hipGraphInstantiate(&ctx->graph_exec, ctx->graph, nullptr, nullptr, 0));
    });
}

int32_t zmpm_graph_launch(zmpm_ctx *ctx) {
    return guarded([&] {
        check_ctx(ctx);
        if (!ctx->graph_exec) throw ZmpmError(ZMPM_ERR_STATE, "zmpm_graph_capture must precede zmpm_graph_launch");
        ctx->set_device();
        // the chain starts on a partition cut as at the capture (and, fused, from the tiles of the
        // current state)
        align_parts(ctx, ctx->graph_shift);
        if (ctx->graph_fused) fused_prologue(ctx, true);
        HIP_CHECK(hipGraphLaunch(ctx->graph_exec, ctx->stream));
        ctx->partitioned = true;
        ctx->grid_ready = false;
        if (ctx->graph_fused) ctx->tiles_valid = true;
    });
}

int32_t zmpm_set_fused(zmpm_ctx *ctx, int32_t on) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->fuse = on != 0;
        ctx->tiles_valid = false;
    });
}

int32_t zmpm_fused(const zmpm_ctx *ctx) { return ctx ? (ctx->fused_path() ? 1 : 0) : -1; }

int64_t zmpm_state_bytes(const zmpm_ctx *ctx) { return ctx ? (int64_t)sizeof(float) * state_floats(ctx->n) : 0; }

int32_t zmpm_state_save(zmpm_ctx *ctx, void *dst) {
    return guarded([&] {
        check_ctx(ctx);
        if (!dst) throw ZmpmError(ZMPM_ERR_ARG, "null buffer");
        ctx->set_device();
        if (!ctx->partitioned) full_partition(ctx);  // saves are block-sorted
        HIP_CHECK(hipMemcpyAsync(dst, ctx->st[ctx->cur].ch, sizeof(float) * state_floats(ctx->n),
                                 hipMemcpyDeviceToDevice, ctx->stream));
    });
}

int32_t zmpm_state_load(zmpm_ctx *ctx, const void *src) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->tiles_valid = false;  // the fused pass's tiles no longer match the state
        if (!src) throw ZmpmError(ZMPM_ERR_ARG, "null buffer");
        ctx->set_device();
        HIP_CHECK(hipMemcpyAsync(ctx->st[ctx->cur].ch, src, sizeof(float) * state_floats(ctx->n),
                                 hipMemcpyDeviceToDevice, ctx->stream));
        ctx->partitioned = false;
        ctx->loaded_sorted = true;
        ctx->grid_ready = false;
    });
}

int32_t zmpm_adjoint_substep(zmpm_ctx *ctx, const zmpm_step_params *p, const zmpm_adjoint *adj) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->tiles_valid = false;  // the fused pass's tiles no longer match the state
        check_step_params(p, 0);
        if (!adj || adj->n != ctx->n) throw ZmpmError(ZMPM_ERR_ARG, "adjoint arrays must cover the resident particles");
        if (ctx->n > 0 && (!adj->x || !adj->v || !adj->C || !adj->F || !adj->mu || !adj->lam))
            throw ZmpmError(ZMPM_ERR_ARG, "null adjoint array");
        if (ctx->has_stvk || ctx->has_aniso)
            throw ZmpmError(ZMPM_ERR_STATE,
                            "the adjoint covers fixed-corotated and J-fluid materials only, without fiber terms");
        if (ctx->L.ncols > 0 || ctx->bnd.n > 0)
            throw ZmpmError(ZMPM_ERR_STATE, "the adjoint covers the sticky box only, not colliders or boundary particles");
        if (ctx->transfer != ZMPM_TRANSFER_APIC) throw ZmpmError(ZMPM_ERR_STATE, "the adjoint covers APIC transfers");
        ctx->set_device();
        if (!ctx->grid_v) {
            ctx->grid_v = (float *)ctx->alloc(sizeof(float) * 256 * ctx->grid_blocks);
            ctx->gbar = (float *)ctx->alloc(sizeof(float) * 256 * ctx->grid_blocks);
            ctx->pay = (float *)ctx->alloc(sizeof(float) * 12 * std::max<int64_t>(ctx->cap, 1));
        }
        if (!ctx->partitioned) full_partition(ctx);
        ctx->L.part_shift = ctx->psize[ctx->pcur];  // the parts the partition was cut into
        apply_step_params(ctx, p);
        hipStream_t s = ctx->stream;
        const State A = ctx->st[ctx->cur];
        const Partition P = ctx->part[ctx->pcur];
        const int64_t npb = ctx->max_units(), ngb = ctx->max_gblocks();
        AdjArrays aa{adj->x, adj->v, adj->C, adj->F, adj->mu, adj->lam, adj->n};
        // forward transfer again: combined mass and momentum in grid, updated velocities in grid_v
        launch_stress(A, ctx->n, ctx->L, ctx->mats, ctx->sig, ctx->cap, s, ctx->stress_general());
        launch_p2g(A, P, ctx->L, ctx->mats, ctx->sig, ctx->cap, ctx->tiles, ctx->status, npb, s, ctx->p2g_sparse,
                   ctx->fused_path());
        launch_grid(GRID_COMBINE | GRID_KEEP | GRID_UPDATE | GRID_BOUNDARY, P, ctx->L, ctx->tiles, ctx->grid,
                    ctx->gvmax, ngb, s, ctx->grid_v);
        // G2P adjoint: particle part, then its velocity scatter into gbar
        const int64_t stride = ctx->cap;
        launch_g2p_adjoint(A, P, ctx->L, ctx->grid_v, aa, ctx->pay, ctx->pay + 3 * stride, stride, npb, s);
        launch_p2g_adjoint(A, P, ctx->L, AdjScatter{ctx->pay, ctx->pay + 3 * stride, stride}, ctx->tiles,
                           ctx->status, npb, s);
        launch_grid(GRID_COMBINE, P, ctx->L, ctx->tiles, ctx->gbar, ctx->gvmax, ngb, s);
        launch_grid_adjoint(P, ctx->L, ctx->grid, ctx->gbar, ngb, s);
        launch_p2g_gather_adjoint(A, P, ctx->L, ctx->mats, ctx->gbar, aa, npb, s);
        HIP_CHECK(hipGetLastError());
        ctx->grid_ready = false;
        ctx->grid_part = ctx->pcur;
    });
}

int64_t zmpm_num_particle_blocks(zmpm_ctx *ctx) {
    if (!ctx) return -1;
    uint32_t c[4] = {0, 0, 0, 0};
    (void)hipSetDevice(ctx->device);
    if (hipMemcpyAsync(c, ctx->part[ctx->pcur].counts, sizeof(c), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        return -1;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return -1;
    return c[0];
}

int64_t zmpm_num_block_parts(zmpm_ctx *ctx) {
    if (!ctx) return -1;
    uint32_t c[4] = {0, 0, 0, 0};
    (void)hipSetDevice(ctx->device);
    if (hipMemcpyAsync(c, ctx->part[ctx->pcur].counts, sizeof(c), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        return -1;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return -1;
    return (int64_t)c[0] + c[2];
}

int64_t zmpm_num_grid_blocks(zmpm_ctx *ctx) {
    if (!ctx) return -1;
    uint32_t c[4] = {0, 0, 0, 0};
    (void)hipSetDevice(ctx->device);
    if (hipMemcpyAsync(c, ctx->part[ctx->grid_part].counts, sizeof(c), hipMemcpyDeviceToHost, ctx->stream) !=
        hipSuccess)
        return -1;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return -1;
    return c[1];
}

int32_t zmpm_grid_download(zmpm_ctx *ctx, int32_t *keys_host, float *data_host, int64_t max_blocks) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->set_device();
        const Partition &P = ctx->part[ctx->grid_part];
        uint32_t c[4];
        HIP_CHECK(hipMemcpyAsync(c, P.counts, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        const int64_t G = c[1];
        if (G > max_blocks) throw ZmpmError(ZMPM_ERR_CAPACITY, "grid download buffer too small");
        std::vector<uint32_t> lat(G), scn(G);
        if (G > 0) {
            HIP_CHECK(hipMemcpyAsync(lat.data(), P.gblk_lat, sizeof(uint32_t) * G, hipMemcpyDeviceToHost, ctx->stream));
            HIP_CHECK(hipMemcpyAsync(scn.data(), P.gblk_scene, sizeof(uint32_t) * G, hipMemcpyDeviceToHost, ctx->stream));
            HIP_CHECK(hipMemcpyAsync(data_host, ctx->grid, sizeof(float) * 256 * G, hipMemcpyDeviceToHost,
                                     ctx->stream));
        }
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        for (int64_t g = 0; g < G; ++g) {
            // block coordinates inside the block's own scene
            const SceneParams &sc = ctx->scenes_host[scn[g] < (uint32_t)ctx->L.nscenes ? scn[g] : 0];
            const int64_t nb = sc.nb;
            const int64_t l = (int64_t)lat[g] - sc.lat_off;
            keys_host[3 * g + 0] = (int32_t)((l / (nb * nb)) + ctx->L.origin) * 4;
            keys_host[3 * g + 1] = (int32_t)(((l / nb) % nb) + ctx->L.origin) * 4;
            keys_host[3 * g + 2] = (int32_t)((l % nb) + ctx->L.origin) * 4;
        }
    });
}

int32_t zmpm_order_download(zmpm_ctx *ctx, int32_t *pid_host, int64_t n) {
    return guarded([&] {
        check_ctx(ctx);
        if (n != ctx->n) throw ZmpmError(ZMPM_ERR_ARG, "n must equal the resident particle count");
        ctx->set_device();
        if (n > 0) {
            launch_extract_pid(ctx->st[ctx->cur], n, ctx->perm, ctx->stream);
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipMemcpyAsync(pid_host, ctx->perm, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->stream));
        }
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int32_t zmpm_bin_keys(zmpm_ctx *ctx, const float *x, int64_t n, uint32_t *keys) {
    return guarded([&] {
        check_ctx(ctx);
        if (n < 0 || (n > 0 && (!x || !keys))) throw ZmpmError(ZMPM_ERR_ARG, "invalid arguments");
        ctx->set_device();
        fill_u32(&ctx->status->domain_error, 0u, 1, ctx->stream);
        launch_bin_keys(x, n, n, ctx->mats, ctx->L, keys, ctx->status, ctx->stream);
        HIP_CHECK(hipGetLastError());
        check_status(ctx);
    });
}

int32_t zmpm_sort_keys(zmpm_ctx *ctx, const uint32_t *keys, int64_t n, int32_t key_bits, uint32_t *keys_out,
                       uint32_t *perm_out) {
    return guarded([&] {
        check_ctx(ctx);
        if (n < 0 || n > ctx->cap) throw ZmpmError(ZMPM_ERR_CAPACITY, "n outside [0, capacity]");
        if (key_bits < 1 || key_bits > 32) throw ZmpmError(ZMPM_ERR_ARG, "key_bits must be in [1, 32]");
        if (n > 0 && (!keys || !keys_out || !perm_out)) throw ZmpmError(ZMPM_ERR_ARG, "null pointer");
        if (n > 0 && (keys_out == keys || (const void *)perm_out == (const void *)keys || (void *)perm_out == (void *)keys_out))
            throw ZmpmError(ZMPM_ERR_ARG, "keys, keys_out and perm_out must be distinct buffers");
        ctx->set_device();
        radix_sort_pairs(keys, nullptr, n, 0, key_bits, keys_out, perm_out, nullptr, ctx->sort_scratch, ctx->stream);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int32_t zmpm_order_by_morton_code(zmpm_ctx *ctx, int32_t layout, int32_t nch, int64_t n, const float *src,
                                  float *dst, int32_t x_chn, const float bv_min[3], const float bv_max[3],
                                  uint32_t *codes, int32_t *ordered_to_original, int32_t *original_to_ordered) {
    return guarded([&] {
        check_ctx(ctx);
        if (n < 0 || n > ctx->cap) throw ZmpmError(ZMPM_ERR_CAPACITY, "n must be in [0, context capacity]");
        if (layout != ZMPM_LAYOUT_SOA && layout != ZMPM_LAYOUT_AOS && layout != ZMPM_LAYOUT_AOSOA32)
            throw ZmpmError(ZMPM_ERR_ARG, "unknown layout");
        if (nch < 3 || x_chn < 0 || x_chn + 3 > nch) throw ZmpmError(ZMPM_ERR_ARG, "x channels outside the buffer");
        if (!bv_min || !bv_max) throw ZmpmError(ZMPM_ERR_ARG, "null bounding box");
        float lo[3], ext[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = bv_min[k];
            ext[k] = bv_max[k] - bv_min[k];
            if (!(ext[k] > 0.0f) || !std::isfinite(ext[k]) || !std::isfinite(lo[k]))
                throw ZmpmError(ZMPM_ERR_ARG, "the bounding box needs a positive finite extent per axis");
        }
        if (n == 0) return;
        if (!src || !dst || !ordered_to_original || !original_to_ordered)
            throw ZmpmError(ZMPM_ERR_ARG, "null buffer");
        if (src == dst) throw ZmpmError(ZMPM_ERR_ARG, "dst must not alias src");
        ctx->set_device();
        hipStream_t s = ctx->stream;
        launch_morton_codes(layout, nch, n, src, x_chn, lo, ext, ctx->keys, s);
        // stable (code, index) sort: radix_sort_pair (Structures.hpp:430-431) and the inverse map (:433-435)
        radix_sort_pairs(ctx->keys, nullptr, n, 0, 30, codes ? codes : ctx->skeys,
                         reinterpret_cast<uint32_t *>(ordered_to_original),
                         reinterpret_cast<uint32_t *>(original_to_ordered), ctx->sort_scratch, s);
        launch_permute_channels(layout, nch, n, src, dst, reinterpret_cast<const uint32_t *>(ordered_to_original), s);
        HIP_CHECK(hipGetLastError());
    });
}

int32_t zmpm_key_bits(const zmpm_ctx *ctx) { return ctx ? ctx->key_bits : -1; }

}  // extern "C"

// ---- allocation-free sort ------------------------------------------------------
// zmpm_sort_keys_noalloc runs the radix sort of zmpm_sort_keys on a scratch that
// lives in the code object: a file-scope device array, which each GPU the library
// runs on holds from the library's load on it to the end of the process.  In host
// code `scratch` only names the symbol; its device address on the current GPU comes
// from hipGetSymbolAddress, looked up once per device and kept in g_noalloc.
namespace {

constexpr int64_t kNoAllocMaxKeys = (int64_t)1 << 20;
constexpr size_t kNoAllocScratchBytes = radix_scratch_bound(kNoAllocMaxKeys);  // 9 MiB + 512 B
constexpr int kNoAllocMaxDevices = 64;

struct NoAllocScratch {
    std::mutex mu;
    void *ptr[kNoAllocMaxDevices] = {};
};
NoAllocScratch g_noalloc;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

// Device address of `scratch` on the current device: on the first call on a device,
// with that device current, check that it is a gfx950 and look the symbol up.
void *noalloc_scratch() {
    int dev = -1;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) throw ZmpmError(ZMPM_ERR_NODEVICE, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_noalloc.mu);
    if (!g_noalloc.ptr[dev]) {
        hipDeviceProp_t prop;
        HIP_CHECK(hipGetDeviceProperties(&prop, dev));
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            throw ZmpmError(ZMPM_ERR_NODEVICE, std::string("libzmpm is built for gfx950, device is ") + prop.gcnArchName);
        void *ptr = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw ZmpmError(ZMPM_ERR_HIP, "hipGetSymbolAddress returned a null address for the sort scratch");
        g_noalloc.ptr[dev] = ptr;
    }
    return g_noalloc.ptr[dev];
}

}  // namespace

extern "C" {

int32_t zmpm_sort_keys_noalloc(const uint32_t *keys, int64_t n, int32_t key_bits, uint32_t *keys_out,
                               uint32_t *perm_out, void *stream) {
    return guarded([&] {
        if (n < 0 || n > kNoAllocMaxKeys)
            throw ZmpmError(ZMPM_ERR_CAPACITY, "n outside [0, 1048576], the static scratch of zmpm_sort_keys_noalloc");
        if (key_bits < 1 || key_bits > 32) throw ZmpmError(ZMPM_ERR_ARG, "key_bits must be in [1, 32]");
        if (n == 0) return;
        if (!keys || !keys_out || !perm_out) throw ZmpmError(ZMPM_ERR_ARG, "null pointer");
        if (keys_out == keys || (const void *)perm_out == (const void *)keys || (void *)perm_out == (void *)keys_out)
            throw ZmpmError(ZMPM_ERR_ARG, "keys, keys_out and perm_out must be distinct buffers");
        void *ptr = noalloc_scratch();
        // kernels only (histograms, scans and scatters rewrite every scratch word they read)
        radix_sort_pairs(keys, nullptr, n, 0, key_bits, keys_out, perm_out, nullptr, ptr, (hipStream_t)stream);
        HIP_CHECK(hipGetLastError());
    });
}

}  // extern "C"
