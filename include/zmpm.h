/* zmpm.h -- C ABI of the MI355X (gfx950) MLS/APIC-MPM substep library (libzmpm.so).
 *
 * Drop-in boundary for Zeno's CuLagrange MPM node chain.  Each entry point
 * below names the reference node (or zpc call site) whose work it replaces:
 *
 *   zmpm_partition    ZSPartitionForZSParticles::apply  projects/CuLagrange/mpm/MPMPipeline.cu:21-112
 *                     + ExpandZSPartition::apply         projects/CuLagrange/mpm/MPMPipeline.cu:121-231
 *   zmpm_grid_clear   ZSGridFromZSPartition::apply      projects/CuLagrange/mpm/MPMPipeline.cu:241-267
 *   zmpm_p2g          ZSParticleToZSGrid::apply (apic)  projects/CuLagrange/mpm/Transfer.cu:985-1059, 844-890
 *   zmpm_grid_update  UpdateZSGrid::apply (apic)        projects/CuLagrange/mpm/MPMPipeline.cu:828-944
 *   zmpm_boundary     ApplyBoundaryOnZSGrid::apply      projects/CuLagrange/mpm/Boundary.cu:83-163
 *   zmpm_g2p          ZSGridToZSParticle::apply (apic)  projects/CuLagrange/mpm/Transfer.cu:1069-1173
 *   zmpm_return_mapping ZSReturnMapping::apply         projects/CuLagrange/mpm/MPMPipeline.cu:953-1145
 *   zmpm_compute_particle_beta ComputeParticleBeta::apply projects/CuLagrange/mpm/Boundary.cu:172-509
 *   zmpm_set_colliders ToZSBoundary::apply             projects/CuLagrange/mpm/Generation.cu:1597-1650
 *   zmpm_step_colliders StepZSBoundary::apply          projects/CuLagrange/mpm/Generation.cu:1663-1672
 *   zmpm_set_boundary_particles ZSBoundaryPrimitiveToZSGrid::apply projects/CuLagrange/mpm/Transfer.cu:1533-1622
 *                     + ApplyGridBoundaryOnZSGrid::apply projects/CuLagrange/mpm/Boundary.cu:18-74
 *   zmpm_max_vel_sqr  UpdateZSGrid "MaxVelSqr" output    projects/CuLagrange/mpm/MPMPipeline.cu:939-942
 *   zmpm_set_materials ConfigConstitutiveModel::apply   projects/CuLagrange/mpm/Generation.cu:18-88
 *   zmpm_import/export ToZSParticles (particle store)    projects/CuLagrange/mpm/Generation.cpp:690-1001
 *                      and the TileVector hand-off       projects/PyZpc/interop/CAPI.cpp:159-180
 *   zmpm_import_objects/export_objects  the "ZSParticles" ListObject (RETRIEVE_OBJECT_PTRS)
 *                                                        projects/CUDA/Utils.hpp:14-26
 *   zmpm_particle_volume ComputeParticleVolume::apply  projects/CuLagrange/mpm/Maintenance.cu:17-56
 *   zmpm_last_error   Zeno_GetLastError / LastError      zeno/include/zeno/extra/CAPI.h:133-163
 *
 * Conventions
 *   - Every function returns a zmpm status (0 = ZMPM_OK).  No C++ exception
 *     crosses this boundary; zmpm_last_error() returns the message of the
 *     last failing call on the calling thread.
 *   - All particle/grid pointers passed in are DEVICE pointers, owned by the
 *     caller.  The library owns its resident particle state, grid, partition
 *     and scratch inside the opaque zmpm_ctx.
 *   - Work is enqueued on the context's HIP stream (zmpm_set_stream; default
 *     stream 0) and is asynchronous, except for calls documented as
 *     synchronising.  A context is not re-entrant; distinct contexts are
 *     independent.
 *   - Arithmetic is FP32; the polar decomposition inside the fixed-corotated
 *     stress runs in FP64.  Binning keys and the particle order are
 *     bit-exact (stable sort by block key).
 */
#ifndef ZMPM_H
#define ZMPM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__) || defined(__clang__)
#define ZMPM_API __attribute__((visibility("default")))
#else
#define ZMPM_API
#endif

typedef struct zmpm_ctx zmpm_ctx;

/* status codes */
#define ZMPM_OK 0
#define ZMPM_ERR_ARG 1      /* invalid argument */
#define ZMPM_ERR_HIP 2      /* HIP runtime error */
#define ZMPM_ERR_CAPACITY 3 /* more particles than the context was created for */
#define ZMPM_ERR_DOMAIN 4   /* a particle left the grid lattice or is not finite */
#define ZMPM_ERR_STATE 5    /* call made out of order (e.g. p2g before partition) */
#define ZMPM_ERR_NODEVICE 6 /* no gfx950 device */

/* constitutive models (Structures.hpp:28-29 ElasticModel; J-fluid = Generation.cpp:878-884 "J" channel) */
#define ZMPM_MODEL_FCR 0          /* fixed corotated, F-based */
#define ZMPM_MODEL_JFLUID 1       /* weakly compressible fluid, J-based: tau = K J (J-1) I */
#define ZMPM_MODEL_STVK_HENCKY 2  /* StVK with Hencky strain, F-based ("stvk", Generation.cu:45-46) */
#define ZMPM_MODEL_NEO_HOOKEAN 3  /* compressible Neo-Hookean, F-based ("nhk", Generation.cu:43-44):
                                     P = mu (F - F^-T) + lam log(J) F^-T (J clamped to >= 1e-6) */
#define ZMPM_MODEL_STABLE_NEO_HOOKEAN 4 /* stable Neo-Hookean, invariant form ("snhk", Generation.cu:47-48):
                                           P = mu F + (lam (J - 1) - mu) cof(F) */

/* plasticity of a ZMPM_MODEL_STVK_HENCKY material (ConfigConstitutiveModel "plasticity",
 * Generation.cu:67-85; the return mapping of ZSReturnMapping, MPMPipeline.cu:953-972) */
#define ZMPM_PLASTIC_NONE 0
#define ZMPM_PLASTIC_DRUCKER_PRAGER 1 /* "nadp": non-associative Drucker-Prager, friction_angle */
#define ZMPM_PLASTIC_VON_MISES 2      /* "navm": non-associative von Mises, yield_stress */
#define ZMPM_PLASTIC_CAM_CLAY 3       /* "nacc": non-associative Cam-clay, friction_angle, nacc_beta, nacc_xi,
                                         hardening on the particle's logJp (Generation.cu:77-82) */

/* anisotropic elastic term added to P ("aniso", ConfigConstitutiveModel, Generation.cu:60-63) */
#define ZMPM_ANISO_NONE 0
#define ZMPM_ANISO_ARAP 1 /* "arap": fiber stiffness aniso_strength * mu(E, nu) along the particle's
                             rest-space fiber "a", the energy of Kim, De Goes, Iben 2019 */

/* grid boundary conditions */
#define ZMPM_BOUNDARY_NONE 0
#define ZMPM_BOUNDARY_STICKY_BOX 1 /* v := 0 on nodes with a coordinate < pad or > n_cells - pad */

/* particle buffer layouts */
#define ZMPM_LAYOUT_SOA 0     /* channel planes: p[c * n + i] */
#define ZMPM_LAYOUT_AOS 1     /* interleaved:    p[i * nch + c] */
#define ZMPM_LAYOUT_AOSOA32 2 /* zpc TileVector<f32,32>: base[((i / 32) * tile_channels + chn + c) * 32 + i % 32] */

typedef struct {
    int32_t model; /* ZMPM_MODEL_* */
    float E;       /* Young's modulus (FCR) */
    float nu;      /* Poisson ratio (FCR) */
    float rho;     /* density: particle mass = rho * vol (Generation.cpp:944-946) */
    float vol;     /* particle volume = dx^3 / ppc (Generation.cu:25) */
    float bulk;    /* bulk modulus K (J-fluid) */
    int32_t scene; /* scene whose particles use this material (0 in a one-scene context) */
    int32_t plasticity;   /* ZMPM_PLASTIC_* (ZMPM_MODEL_STVK_HENCKY only) */
    float friction_angle; /* degrees, Drucker-Prager (reference default 35) */
    float yield_stress;   /* von Mises yield stress on |dev tau| (reference default 1e5) */
    float flip_beta;      /* ZMPM_TRANSFER_AFLIP positional adjustment beta (0: AFLIP, > 0: ASFLIP): the
                             initial value of each particle's beta channel */
    float nacc_beta;      /* Cam-clay cohesion ratio beta (reference default 2) */
    float nacc_xi;        /* Cam-clay hardening xi (reference default 1) */
    int32_t aniso;        /* ZMPM_ANISO_* (F-based models; Transfer.cu:866-867) */
    float aniso_strength; /* ARAP fiber strength (reference default 10, Generation.cu:62) */
} zmpm_material;

/* particle-grid transfer schemes (ZenoGrid transferScheme, Structures.hpp:606-618) */
#define ZMPM_TRANSFER_APIC 0  /* "apic" */
#define ZMPM_TRANSFER_AFLIP 1 /* "aflip" / "asflip": momentum and stress scattered apart (Transfer.cu:935-984),
                                 v = vstar + 0.99 (v_p - v), x += dt (vstar + beta 0.99 (v_p - v))
                                 with the particle's beta channel (Transfer.cu:1175-1214) */
#define ZMPM_TRANSFER_FLIP 2  /* "flip": no affine term; momentum W m v, stress W (-dt vol P F^T grad W)
                                 (Transfer.cu:892-933); v = vstar + 0.99 (v_p - v), x += dt vstar,
                                 F <- (I + dt sum vstar grad W^T) F, C kept (Transfer.cu:1215-1250) */

typedef struct {
    float dx;         /* grid spacing (MakeZSGrid "dx", Generation.cu:1486-1519) */
    int32_t n_cells;  /* domain side in cells; the block lattice covers [-4, 4*ceil(n_cells/4)+8)
                         (all scenes together: fewer than 2^31 blocks; 26 bytes of lattice maps per
                         block, everything else sized by the capacity) */
    int64_t capacity; /* maximum number of particles */
    int32_t device;   /* HIP device ordinal */
} zmpm_config;

/* One independent scene of a multi-scene context: its own grid spacing,
 * domain, time step, gravity and walls (one ZenoPartition/ZenoGrid pair,
 * Structures.hpp:533-623, with the parameters of its substep nodes). */
typedef struct {
    float dx;
    int32_t n_cells;
    float dt;
    float gravity[3];
    int32_t boundary; /* ZMPM_BOUNDARY_* */
    int32_t pad;      /* sticky wall thickness in cells */
} zmpm_scene;

typedef struct {
    float dt;
    float gravity[3];  /* UpdateZSGrid gravity / Accel (MPMPipeline.cu:837-846) */
    int32_t boundary;  /* ZMPM_BOUNDARY_* */
    int32_t pad;       /* sticky wall thickness in cells */
} zmpm_step_params;

/* Level-set colliders (ToZSBoundary, projects/CuLagrange/mpm/Generation.cu:1597-1650; zpc's
 * Collider, absent).  A grid node with m != 0 inside the object (phi < 0) gets the response of
 * ApplyGridBoundaryOnZSGrid (Boundary.cu:55-66) with the object velocity
 * v_obj = translation_rate + angular_velocity x (x - b) + (scale_rate / scale) (x - b):
 * sticky v = v_obj; slip removes the normal component of v - v_obj; separate removes it only
 * when v - v_obj points into the object.  The normal is the rotated gradient of phi. */
#define ZMPM_COLLIDER_STICKY 0
#define ZMPM_COLLIDER_SLIP 1
#define ZMPM_COLLIDER_SEPARATE 2

/* shapes, phi in the collider frame X = R^T (x - b) / scale, distances scaled back by scale */
#define ZMPM_SHAPE_HALFSPACE 0 /* phi = n . X - d: shape_params[0..2] = unit n, [3] = d */
#define ZMPM_SHAPE_SPHERE 1    /* phi = |X - c| - r: [0..2] = c, [3] = r */
#define ZMPM_SHAPE_BOX 2       /* exact box distance: [0..2] = lo, [3..5] = hi */
#define ZMPM_SHAPE_SDF_GRID 3  /* trilinear in samples sdf[(i * dims[1] + j) * dims[2] + k] at
                                  origin + spacing (i, j, k); outside the sample box phi = +inf */
/* sparse 8^3-sample leaves (OpenVDB's leaf nodes; the SparseGrid that ToZSLevelSet builds from a
 * VDB FloatGrid, Generation.cpp:1536-1547): leaf l holds sdf[l * 512 + (i & 7) * 64 + (j & 7) * 8
 * + (k & 7)] for the samples (i, j, k) of the 8^3 box starting at leaf_origins[3 l .. 3 l + 2]
 * (multiples of 8, distinct); samples at origin + spacing (i, j, k), trilinear as SDF_GRID; a
 * sample outside every leaf reads the background shape_params[0] (a VDB level set's outside
 * value).  zmpm_set_colliders reads the origins once and keeps a dense map over the leaves'
 * bounding box (4 bytes per 8^3 box of it, held until zmpm_destroy); the samples are read at
 * every substep. */
#define ZMPM_SHAPE_SDF_LEAVES 4

typedef struct {
    int32_t type;           /* ZMPM_COLLIDER_* */
    int32_t shape;          /* ZMPM_SHAPE_* */
    int32_t scene;          /* scene it acts in (0 in a one-scene context) */
    int32_t dims[3];        /* SDF_GRID sample counts (>= 2 each) */
    float shape_params[6];
    const float *sdf;       /* SDF_GRID / SDF_LEAVES device samples, caller-owned, read at every substep */
    float origin[3];
    float spacing;
    float translation[3];   /* b */
    float translation_rate[3];
    float scale;            /* s > 0 */
    float scale_rate;
    float ypr_degrees[3];   /* R = Ry(yaw) Rx(pitch) Rz(roll) */
    float angular_velocity[3];
    const int32_t *leaf_origins; /* SDF_LEAVES: device (nleaves, 3) index coordinates, caller-owned */
    int32_t nleaves;        /* SDF_LEAVES leaf count (>= 1); sdf then holds nleaves * 512 samples */
    int32_t reserved;       /* 0 */
} zmpm_collider;

/* A view of caller-owned particle attribute buffers (device pointers).
 * For ZMPM_LAYOUT_SOA / AOS each attribute has its own base pointer and the
 * *_chn fields are ignored.  For ZMPM_LAYOUT_AOSOA32 all pointers are the
 * same TileVector base and *_chn is the attribute's first channel
 * (e.g. the "mpm" FCR map of Generation.cpp:868-884: x=1, v=4, C=8, F=19).
 * F (9 channels, row-major) is read for FCR particles, J (1 channel) for
 * J-fluid particles; either may be NULL when no particle uses it. */
typedef struct {
    int64_t n;
    int32_t layout;
    int32_t tile_channels;
    float *x;
    int32_t x_chn;
    float *v;
    int32_t v_chn;
    float *C;
    int32_t C_chn;
    float *F;
    int32_t F_chn;
    float *J;
    int32_t J_chn;
    const int32_t *material; /* per-particle index into the material table (always a plain int32 array) */
    float *beta;  /* optional (NULL: the material's flip_beta on import, not written on export): */
    int32_t beta_chn; /*   ASFLIP positional adjustment, the "beta" channel (Generation.cpp:989) */
    float *logJp; /* optional (NULL: -0.04 on import, Generation.cpp:987, not written on export): */
    int32_t logJp_chn; /*  Cam-clay plastic volume log, the "logJp" channel (Generation.cpp:899) */
    float *mu;    /* optional, with lam: per-particle Lame parameters ("mu", "lam" channels) that */
    int32_t mu_chn; /*    replace the material's in the stress (Transfer.cu:861-864; not under FLIP, */
    float *lam;   /*      whose P2G has no override, Transfer.cu:892-910).  Both NULL on import: */
    int32_t lam_chn; /*   the material's; exported as the particle's values */
    float *a;     /* optional: the rest-space fiber direction of ZMPM_ANISO_ARAP, the "a" channel */
    int32_t a_chn; /*   (3 channels, Generation.cpp:892,977-982); NULL on import: (0, 1, 0) */
    float *m;     /* optional, with vol: per-particle mass and volume, the "m" and "vol" channels */
    int32_t m_chn; /*   (0 and 7 in the "mpm" map, Generation.cpp:868-884) that P2G reads per particle */
    float *vol;   /*    (Transfer.cu:857-858) and ComputeParticleVolume writes (Maintenance.cu:50-51). */
    int32_t vol_chn; /* Both NULL on import: the material's rho * vol and vol; exported as the particle's */
    int32_t default_material; /* material of every particle when material is NULL (one ZenoParticles
                                 object has one model, Structures.hpp:134-186) */
} zmpm_particles;

/* ---- context ---------------------------------------------------------------- */
ZMPM_API int32_t zmpm_create(zmpm_ctx **out, const zmpm_config *cfg);
/* A context holding nscenes independent scenes (no counterpart in the reference, which runs
 * one scene per node chain).  Each scene owns a contiguous part of one block lattice, and every
 * kernel launch of a substep covers all scenes; a particle belongs to the scene of its material
 * (zmpm_material.scene).  zmpm_create is this with one scene. */
ZMPM_API int32_t zmpm_create_scenes(zmpm_ctx **out, const zmpm_scene *scenes, int32_t nscenes, int64_t capacity,
                                    int32_t device);
ZMPM_API int32_t zmpm_num_scenes(const zmpm_ctx *ctx);
ZMPM_API int32_t zmpm_destroy(zmpm_ctx *ctx);
ZMPM_API const char *zmpm_last_error(void);
ZMPM_API int32_t zmpm_set_stream(zmpm_ctx *ctx, void *hip_stream);
ZMPM_API int32_t zmpm_sync(zmpm_ctx *ctx); /* synchronising; reports deferred device errors (ZMPM_ERR_DOMAIN) */

/* ConfigConstitutiveModel (Generation.cu:18-88): E, nu -> Lame parameters, volume, density. */
ZMPM_API int32_t zmpm_set_materials(zmpm_ctx *ctx, const zmpm_material *mats, int32_t count);

/* ToZSParticles (Generation.cpp:690-1001) / TileVector hand-off (CAPI.cpp:159-180):
 * copy caller particles into the resident state and bin them (stable sort by block key). */
ZMPM_API int32_t zmpm_import(zmpm_ctx *ctx, const zmpm_particles *view);
/* write the resident state back into caller buffers, in the caller's particle order */
ZMPM_API int32_t zmpm_export(zmpm_ctx *ctx, const zmpm_particles *view);
/* The "ZSParticles" input as a list of particle objects (RETRIEVE_OBJECT_PTRS, projects/CUDA/Utils.hpp:14-26;
 * the per-object loops of Transfer.cu:997-1055): nviews views, one per ZenoParticles object, each with its
 * own layout, tile_channels and channel offsets (e.g. a J-fluid TileVector of 20 channels with J at 19
 * and a fixed-corotated one of 28 with F at 19, Generation.cpp:868-884), read in place with no
 * concatenation copy.  Particle i of view k gets the caller index n_0 + ... + n_{k-1} + i. */
ZMPM_API int32_t zmpm_import_objects(zmpm_ctx *ctx, const zmpm_particles *views, int32_t nviews);
/* write the resident state back into the objects' buffers (views as given to zmpm_import_objects) */
ZMPM_API int32_t zmpm_export_objects(zmpm_ctx *ctx, const zmpm_particles *views, int32_t nviews);
/* Host-side check of a view (no device access, no context): a known layout, n >= 0, x present, and for
 * ZMPM_LAYOUT_AOSOA32 tile_channels > 0 and every present attribute's channels inside the tile
 * (chn >= 0, chn + width <= tile_channels; widths x, v, a 3; C, F 9; J, beta, logJp, mu, lam, m, vol 1).
 * zmpm_import / zmpm_export and their _objects forms run it on every view; ZMPM_ERR_ARG if it fails. */
ZMPM_API int32_t zmpm_check_particles(const zmpm_particles *view);
ZMPM_API int64_t zmpm_num_particles(const zmpm_ctx *ctx);

/* ComputeParticleVolume (projects/CuLagrange/mpm/Maintenance.cu:17-56, counts from
 * spatial_hashing, projects/CUDA/Utils.hpp:770-843): for every particle,
 * vol = dx^3 / (particles of its scene in its cell) and m = rho * vol, written in the
 * caller's particle order to one channel each of caller device buffers in a ZMPM_LAYOUT_*
 * (for ZMPM_LAYOUT_AOSOA32, vol and mass are the TileVector base and *_chn the "vol" and
 * "m" channels, 7 and 0 in the "mpm" map of Generation.cpp:868-884); either may be NULL.
 * The cell of x is floor(x / dx) on each axis.  As in the reference, where P2G reads the
 * particles' "m" and "vol" (Transfer.cu:857-858), the resident particles keep these values as
 * their own mass and volume from then on (exported through zmpm_particles.m / .vol); the
 * materials are not changed.  The grid is used as scratch, so a zmpm_p2g must follow before the
 * next node-chain grid stage. */
ZMPM_API int32_t zmpm_particle_volume(zmpm_ctx *ctx, int32_t layout, int32_t tile_channels, float *vol,
                                      int32_t vol_chn, float *mass, int32_t mass_chn);

/* ---- node-granular stages (the reference node chain, one call per node) ------ */
/* ZSPartitionForZSParticles (MPMPipeline.cu:21-112) + ExpandZSPartition (:121-231):
 * block keys of every particle, stable sort by key, particle blocks B and grid blocks B + {0,1}^3. */
ZMPM_API int32_t zmpm_partition(zmpm_ctx *ctx);
/* ZSGridFromZSPartition (MPMPipeline.cu:241-267): zero the grid (m, v) of the partition's blocks. */
ZMPM_API int32_t zmpm_grid_clear(zmpm_ctx *ctx);
/* ZSParticleToZSGrid, apic (Transfer.cu:844-890): grid m += W m_p; grid v (momentum) += W (m_p (v_p + C_p d) - dt D^-1 vol P F^T d). */
ZMPM_API int32_t zmpm_p2g(zmpm_ctx *ctx, float dt);
/* UpdateZSGrid, apic (MPMPipeline.cu:852-877): v = v / m + a dt on nodes with m != 0; keeps max |v|^2 on the device. */
ZMPM_API int32_t zmpm_grid_update(zmpm_ctx *ctx, float dt, const float accel[3]);
/* The context's transfer scheme (ZMPM_TRANSFER_*; APIC when never called).  With AFLIP or FLIP
 * the boundary acts on vstar (Boundary.cu:93-103), the maximum speed covers |v| and |vstar|, and
 * the particle blocks all take the workgroup P2G.  Switching from APIC to AFLIP or FLIP resets
 * every resident particle's beta to its material's flip_beta. */
ZMPM_API int32_t zmpm_set_transfer(zmpm_ctx *ctx, int32_t scheme);
/* ApplyBoundaryOnZSGrid (Boundary.cu:104-116) with an analytic sticky box collider, then the
 * context's level-set colliders (zmpm_set_colliders), then its boundary particles
 * (zmpm_set_boundary_particles, ApplyGridBoundaryOnZSGrid). */
ZMPM_API int32_t zmpm_boundary(zmpm_ctx *ctx, int32_t boundary, int32_t pad);
/* ApplyBoundaryOnZSGrid's ZSBoundary inputs (ToZSBoundary, Generation.cu:1597-1650): replace the
 * context's colliders (at most 64; count 0 removes them).  zmpm_substeps applies them after the
 * grid update of every substep and then advances each by its scene's dt (StepZSBoundary,
 * Generation.cu:1663-1672: translation += translation_rate dt, scale += scale_rate dt).  With
 * colliders, the per-substep maximum speed includes post-collision node velocities.  A graph
 * captured before this call keeps the colliders it was captured with. */
ZMPM_API int32_t zmpm_set_colliders(zmpm_ctx *ctx, const zmpm_collider *cols, int32_t count);
/* Boundary particles (a boundary primitive sampled as particles: "x", "v", "m", "nrm" channels,
 * Transfer.cu:1546-1549), caller-owned device buffers read at every substep. */
typedef struct {
    int64_t n;
    const float *x;   /* (3, n) SoA positions */
    const float *v;   /* (3, n) SoA velocities */
    const float *m;   /* (n) masses */
    const float *nrm; /* (3, n) SoA normals; may be NULL for ZMPM_COLLIDER_STICKY */
    int32_t scene;    /* scene they bound (0 in a one-scene context) */
} zmpm_boundary_particles;
/* ZSBoundaryPrimitiveToZSGrid (Transfer.cu:1533-1622) + ApplyGridBoundaryOnZSGrid
 * (Boundary.cu:18-74): every substep, after the grid update and the colliders, the boundary
 * particles are transferred with the linear kernel (base floor(x / dx), weights 1 - f, f) into
 * m, m v and m nrm on the simulation grid's nodes (the boundary grid's "boundary" update,
 * MPMPipeline.cu:896-910: v_b = m v / m, n_b = normalize(m nrm / m)), and every node with mass in
 * both grids gets type's response (ZMPM_COLLIDER_*): sticky v = v_b, slip / separate remove the
 * normal part of v - v_b along n_b (separate only when approaching).  The sums are 64-bit
 * fixed point (scaled by a power of two from the largest |m| max(1, |v|, |nrm|)), so the transfer is
 * deterministic.  Count 0 (bp NULL or n 0) removes the boundary.  A graph captured before this
 * call keeps the boundary it was captured with. */
ZMPM_API int32_t zmpm_set_boundary_particles(zmpm_ctx *ctx, const zmpm_boundary_particles *bp, int32_t type);
/* StepZSBoundary (Generation.cu:1663-1672) for the node-by-node chain. */
ZMPM_API int32_t zmpm_step_colliders(zmpm_ctx *ctx, float dt);
/* ZSGridToZSParticle, apic (Transfer.cu:1137-1173): v, C from the grid; x += v dt; F <- (I + dt C) F (J-fluid: J <- (1 + dt tr C) J). */
ZMPM_API int32_t zmpm_g2p(zmpm_ctx *ctx, float dt);
/* ZSReturnMapping::apply (projects/CuLagrange/mpm/MPMPipeline.cu:953-1145, the "mpm" branch
 * :958-972): project the F of every particle whose material has plasticity (Drucker-Prager or
 * von Mises on the StVK-Hencky strain).  zmpm_substeps does this after every G2P itself. */
ZMPM_API int32_t zmpm_return_mapping(zmpm_ctx *ctx);
/* ComputeParticleBeta::apply with a ZSBoundary level set (Boundary.cu:284-313, 398-440): for every
 * particle, at x + dt v: inside a collider of its scene (phi < 0) and moving into it,
 * (v - dbdt) . (R n) < 0, gives beta = 0; otherwise beta = 0.5 if det F (J for J-fluid) >
 * jp_critical and 0 if not; jp_critical < 0 takes the reference's JpcDefault, 1 for every
 * material (its 1.1 branch for Cam-clay, Boundary.cu:389-391, tests the anisotropic-model
 * variant of CUDA/Structures.hpp:30, which never holds Cam-clay, so it never runs).  Writes the particles' beta channel,
 * which the AFLIP transfer reads; dt is the scene's (p->dt, or each scene's own when p is NULL). */
ZMPM_API int32_t zmpm_compute_particle_beta(zmpm_ctx *ctx, const zmpm_step_params *p, float jp_critical);
/* UpdateZSGrid "MaxVelSqr" output (MPMPipeline.cu:939-942), the maximum over all scenes; synchronising. */
ZMPM_API int32_t zmpm_max_vel_sqr(zmpm_ctx *ctx, float *out);
/* The device address of the per-scene MaxVelSqr that every fused substep (zmpm_substeps, graph
 * replays) writes after its grid update, as UpdateZSGrid emits it every substep (MPMPipeline.cu:
 * 939-942): nscenes floats, stream ordered on the context stream, readable with no host
 * synchronisation (e.g. by a kernel computing the next CFL time step).  The address stays valid
 * for the context's lifetime. */
ZMPM_API int32_t zmpm_max_vel_sqr_device(zmpm_ctx *ctx, const float **out);
/* The same per scene: out[s] for s < min(nscenes, number of scenes); synchronising. */
ZMPM_API int32_t zmpm_scene_max_vel_sqr(zmpm_ctx *ctx, float *out, int32_t nscenes);

/* ---- fused substep chain ------------------------------------------------------ */
/* nsteps full substeps (partition, P2G, grid update, boundary, G2P) on the resident state.
 * Deterministic; no host synchronisation inside.  params applies to every scene; NULL runs each
 * scene with its own zmpm_scene parameters. */
ZMPM_API int32_t zmpm_substeps(zmpm_ctx *ctx, const zmpm_step_params *params, int32_t nsteps);

/* Time each candidate workgroup size of the two G2P kernels (64 or 256 threads per particle
 * block), P2G with and without the one-wave kernel for blocks of at most 64 particles, and (APIC)
 * the fused G2P / P2G pass, over nsteps substeps from the current state, keep the fastest, and
 * restore the state (no counterpart in the reference).  Results do not depend on the choice.
 * choice (optional, 3 entries) receives the G2P move size, the G2P update size (0 when the fused
 * pass won: it replaces G2P update, the stress and P2G kernels) and 1 if the one-wave P2G is on. */
ZMPM_API int32_t zmpm_autotune(zmpm_ctx *ctx, const zmpm_step_params *params, int32_t nsteps, int32_t *choice);

/* The fused pass on (default) or off for the APIC transfer (no counterpart in the reference):
 * G2P of one substep and the stress and P2G of the next in one kernel per block of the new
 * partition, whose particles are read once and written once per substep; v and C are stored by
 * the last substep of each zmpm_substeps call (and of a captured graph), since nothing reads them
 * in between.  Results are bit-identical either way.  zmpm_fused: 1 if the fused pass runs. */
ZMPM_API int32_t zmpm_set_fused(zmpm_ctx *ctx, int32_t on);
ZMPM_API int32_t zmpm_fused(const zmpm_ctx *ctx);

/* Capture nsteps (even) fused substeps into a HIP graph on the context stream, replayed by
 * zmpm_graph_launch (no counterpart in the reference).  The state must not be re-imported or
 * re-partitioned between capture and launch; capturing again replaces the graph. */
ZMPM_API int32_t zmpm_graph_capture(zmpm_ctx *ctx, const zmpm_step_params *params, int32_t nsteps);
ZMPM_API int32_t zmpm_graph_launch(zmpm_ctx *ctx);

/* zmpm_substeps with HIP events recorded on the context stream around each phase of every substep
 * (the per-node timers of zeno's ZENO_BENCHMARKING build, zeno/src/core/INode.cpp:12-14,141-143).
 * phase_ms[k] receives the GPU milliseconds of phase k summed over the substeps; synchronising.
 * Phases (one kernel launch each): 0 stress (fixed corotated, FP64 polar), 1 P2G, 2 grid (tile
 * combine, update, boundary, CFL max), 3 G2P move (direction of each particle's new block),
 * 4 partition (lattice scan: next block list), 5 G2P update written in the next stable block order. */
ZMPM_API int32_t zmpm_substeps_timed(zmpm_ctx *ctx, const zmpm_step_params *params, int32_t nsteps,
                                     float *phase_ms, int32_t nphases);

/* ---- differentiation ------------------------------------------------------------
 * No counterpart in the reference (its MPM nodes have no adjoint).  One substep is
 * differentiated at a time, on the resident state at the substep's start: callers
 * checkpoint states with zmpm_state_save / zmpm_state_load and run the adjoint
 * substeps in reverse order (zeno_amd/diff.py does this, every k substeps). */
typedef struct {
    int64_t n;  /* resident particle count */
    /* device planes indexed by the caller's particle index: x (3, n), v (3, n), C (9, n), F (9, n)
     * (J-fluid particles: F[0] is dL/dJ).  In: adjoints of the substep's outputs.  Out: adjoints of
     * its inputs (overwritten). */
    float *x, *v, *C, *F;
    /* (n): per fixed-corotated particle, dL/dmu and dL/dlambda of this substep are added */
    float *mu, *lam;
} zmpm_adjoint;

/* bytes of a saved resident state (33 floats per particle, padded to 64 particles) */
ZMPM_API int64_t zmpm_state_bytes(const zmpm_ctx *ctx);
/* copy the resident state to / from a caller device buffer of zmpm_state_bytes bytes.  A save
 * bins the state first if it is not binned, so saves are in stable block order; a load re-bins
 * without sorting before the next substep.  A loaded state must come from a save of this
 * context (one that is not in block order is reported as ZMPM_ERR_STATE at the next sync). */
ZMPM_API int32_t zmpm_state_save(zmpm_ctx *ctx, void *dst);
ZMPM_API int32_t zmpm_state_load(zmpm_ctx *ctx, const void *src);
/* Reverse-mode substep: the resident state is the substep's input (left unchanged); adj holds the
 * adjoints of its outputs on entry and of its inputs on exit.  params as for zmpm_substeps. */
ZMPM_API int32_t zmpm_adjoint_substep(zmpm_ctx *ctx, const zmpm_step_params *params, const zmpm_adjoint *adj);

/* ---- introspection (synchronising; for tests and tools) ------------------------ */
ZMPM_API int64_t zmpm_num_particle_blocks(zmpm_ctx *ctx);
ZMPM_API int64_t zmpm_num_grid_blocks(zmpm_ctx *ctx);
/* Work units of the current partition: a particle block of d particles is cut into ceil(d / 1024)
 * parts (the last takes the rest), each its own P2G and G2P workgroup (one part per block while
 * the fused G2P / P2G pass is on).  No counterpart in the reference. */
ZMPM_API int64_t zmpm_num_block_parts(zmpm_ctx *ctx);
/* Host copies of the grid last written (by zmpm_p2g / zmpm_grid_* or the last fused substep):
 * keys[G*3] (block origin in cells), data[G*4*64] (m, v0, v1, v2; cell = (lx*4+ly)*4+lz). */
ZMPM_API int32_t zmpm_grid_download(zmpm_ctx *ctx, int32_t *keys_host, float *data_host, int64_t max_blocks);
/* Host copy of the resident particle order: pid[k] = caller index of the particle at sorted position k. */
ZMPM_API int32_t zmpm_order_download(zmpm_ctx *ctx, int32_t *pid_host, int64_t n);

/* ---- binning primitives (bit-exact parity) ------------------------------------ */
/* Block key of each particle (MPMPipeline.cu:75-84: coord = floor(x * (1/dx) - 0.5), block = coord - (coord & 3)),
 * encoded as (lattice block index << 6) | cell id, or, for lattices of 2^26 blocks or more, the lattice
 * block index alone (zmpm_key_bits then counts no cell bits).  x: SoA device planes (3 * n).  keys:
 * device (n). */
ZMPM_API int32_t zmpm_bin_keys(zmpm_ctx *ctx, const float *x, int64_t n, uint32_t *keys);
/* Stable LSD radix sort of (key, index) pairs by the low key_bits bits: keys_out = sorted keys,
 * perm_out[k] = input index of the k-th element (== std::stable_sort by key).  Device pointers. */
ZMPM_API int32_t zmpm_sort_keys(zmpm_ctx *ctx, const uint32_t *keys, int64_t n, int32_t key_bits, uint32_t *keys_out,
                       uint32_t *perm_out);
/* number of significant key bits of this context's lattice (6 cell bits, below 2^26 blocks, + lattice index bits) */
ZMPM_API int32_t zmpm_key_bits(const zmpm_ctx *ctx);

/* ZenoParticles::orderByMortonCode (projects/CUDA/Structures.hpp:406-443) on a caller buffer of
 * n particles with nch float channels (ZMPM_LAYOUT_SOA: channel c of particle i at c * n + i; AOS:
 * i * nch + c; AOSOA32: the TileVector with tile_channels = nch), positions at channels
 * x_chn .. x_chn + 2: the u32 Morton code of each position in the box's uniform coordinates
 * (AABBBox::getUniformCoord: (x - min) / (max - min); per axis trunc(min(max(c * 1024, 0), 1023)),
 * bits interleaved x y z from the high bit), a stable sort of (code, index) (radix_sort_pair),
 * ordered_to_original[k] and its inverse original_to_ordered[i] (vertMapping), and every channel
 * of src permuted into dst in code order (dst particle k = src particle ordered_to_original[k]).
 * Device pointers on the context's device; dst must not alias src; codes (optional) receives
 * the sorted codes.  n at most the context capacity (the sort uses the context's buffers).
 * The reference's element ordering (IncludeElement) belongs to mesh primitives, not particles. */
ZMPM_API int32_t zmpm_order_by_morton_code(zmpm_ctx *ctx, int32_t layout, int32_t nch, int64_t n, const float *src,
                                           float *dst, int32_t x_chn, const float bv_min[3], const float bv_max[3],
                                           uint32_t *codes, int32_t *ordered_to_original,
                                           int32_t *original_to_ordered);

/* zmpm_sort_keys without a context and without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, hosts that forbid hipMalloc).  The sort's scratch is a static 9 MiB
 * device array inside libzmpm.so, held on every GPU the library runs on until the process ends,
 * which bounds n to 1,048,576 keys (ZMPM_ERR_CAPACITY above that).  Runs on the current device,
 * asynchronously on stream (0 = default stream); same results as zmpm_sort_keys (stable, bit-exact).
 * keys, keys_out and perm_out must be distinct.  Every call on one GPU shares that scratch: calls
 * that overlap on one GPU (other streams or threads) race and must be serialised by the caller. */
ZMPM_API int32_t zmpm_sort_keys_noalloc(const uint32_t *keys, int64_t n, int32_t key_bits, uint32_t *keys_out,
                                        uint32_t *perm_out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* ZMPM_H */
