// Adjoint (reverse-mode) kernels of one MLS/APIC-MPM substep for gfx950.
//
// The forward substep (P2G Transfer.cu:844-890, grid update MPMPipeline.cu:852-877,
// sticky box Boundary.cu:104-116, G2P Transfer.cu:1137-1173) is differentiated by
// hand.  The adjoint of a scatter is a gather and the adjoint of a gather is a
// scatter, so the same block decomposition serves both directions:
//   g2p adjoint  one workgroup per particle block, grid velocities staged in LDS:
//                per particle the adjoint of x' = x + dt v', F' = (I + dt C') F
//                (J' = (1 + dt tr C') J) and of the weights' dependence on x; the
//                velocity-adjoint payload (a, B) for the scatter (P2G kernel in
//                adjoint mode, deterministic tiles).
//   grid adjoint one wave per grid block: v = mv / m + dt g and the sticky box.
//   p2g adjoint  one workgroup per particle block, (dL/dm, dL/dmv) staged in LDS:
//                per particle the adjoint of m_i += W m, mv_i += W (m v + B d) with
//                B = m C - dt D^-1 vol P F^T, and of the stress P(F) (FP64 polar).
// Adjoint arrays are indexed by the caller's particle index (the resident
// state carries it), so they stay put while the state is re-binned.
#include "zmpm_internal.h"
#include "mpm_math.h"
#include "mpm_device.h"
#include "../../include/zmpm.h"

namespace zmpm {

namespace {

// The two adjoint gathers treat every particle on its own, so a particle block is
// split over kAdjChunks workgroups of kAdjThreads (each stages the block's grid
// values): small differentiable scenes have too few blocks to fill the GPU.
constexpr int kAdjThreads = 64;
constexpr int kAdjChunks = 8;

// d w / d f of the quadratic B-spline weights of bspline_w
__device__ __forceinline__ void bspline_dw(float f, float dw[3]) {
    dw[0] = f - 1.5f;
    dw[1] = -2.0f * (f - 1.0f);
    dw[2] = f - 0.5f;
}

// Per grid block, dL/dm and dL/dmv of the 6^3 stencil region as float4 (m, mv).
__device__ __forceinline__ void stage_block_grid4(float4 *s_g, const float *grid, const int32_t *lat_gid, int64_t l,
                                                  int64_t nb, int64_t nb2) {
    for (int k = threadIdx.x; k < kTileNodes; k += kAdjThreads) {
        const int tx = k / 36, ty = (k / 6) % 6, tz = k % 6;
        const int ex = tx >> 2, ey = ty >> 2, ez = tz >> 2;
        const int32_t g = lat_gid[l + ex * nb2 + ey * nb + ez];
        const int cell = ((tx & 3) * 4 + (ty & 3)) * 4 + (tz & 3);
        const float *gv = grid + (int64_t)(g >= 0 ? g : 0) * 256 + cell;
        const float v0 = gv[0], v1 = gv[64], v2 = gv[128], v3 = gv[192];
        s_g[k] = g >= 0 ? make_float4(v0, v1, v2, v3) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// G2P adjoint.  In: the resident state (the substep's input), the updated grid
// velocities, and in adj the adjoints of the substep's outputs x', v', C', F'|J'.
// Out: adj.x and adj.F hold the G2P part of the inputs' adjoints (the P2G
// adjoint adds the rest); pay receives, per particle in resident order, the
// scatter payload a = dL/dv' + dt dL/dx' and B = D^-1 (dL/dC' + dt dL/dF' F^T).
__global__ void __launch_bounds__(kAdjThreads) k_g2p_adjoint(State A, Partition P, Lattice L, const float *gridv,
                                                          AdjArrays adj, float *pay_a, float *pay_B, int64_t stride) {
    __shared__ float4 s_v[kTileNodes];
    const int t = threadIdx.x;
    const uint32_t nblocks = P.counts[0];
    const int64_t n = adj.n;
    const uint32_t chunk = blockIdx.x % kAdjChunks;
    for (uint32_t b = blockIdx.x / kAdjChunks; b < nblocks; b += gridDim.x / kAdjChunks) {
        const int64_t l = P.pblk_lat[b];
        const SceneParams sc = scene_params(L, L.nscenes == 1 ? 0u : P.pblk_scene[b]);
        const int64_t nb = sc.nb, nb2 = nb * nb;
        int lb[3];
        lattice_coords(sc, l, lb[0], lb[1], lb[2]);
        const int bc[3] = {(lb[0] + L.origin) * 4, (lb[1] + L.origin) * 4, (lb[2] + L.origin) * 4};
        stage_block_velocity<kAdjThreads>(s_v, gridv, P.lat_gid, l, nb, nb2);
        __syncthreads();
        const uint32_t ps = P.pblk_start[b], pe = P.pblk_start[b + 1];
        for (uint32_t i = ps + chunk * kAdjThreads + t; i < pe; i += kAdjChunks * kAdjThreads) {
            float xp[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) xp[k] = chan(A.ch, i, kChX + k);
            const uint32_t mw = chan_u(A.ch, i, kChMat), pid = chan_u(A.ch, i, kChPid);
            const bool fcr = (mw & kMatF) != 0u;
            Arena ar;
            make_arena(xp, sc.dxinv, ar);
            const int lc[3] = {clamp03(ar.base[0] - bc[0]), clamp03(ar.base[1] - bc[1]), clamp03(ar.base[2] - bc[2])};
            float vel[3], Cn[3][3];
            g2p_gather<true>(s_v, ar, lc, sc.dx, sc.Dinv, vel, Cn);
            float xb[3], vb[3], Cb[3][3], Fb[3][3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                xb[k] = adj.x[k * n + pid];
                vb[k] = adj.v[k * n + pid] + sc.dt * xb[k];  // x' = x + dt v'
            }
#pragma unroll
            for (int q = 0; q < 9; ++q) Cb[q / 3][q % 3] = adj.C[q * n + pid];
            if (fcr) {
                float F[3][3];
#pragma unroll
                for (int q = 0; q < 9; ++q) {
                    F[q / 3][q % 3] = chan(A.ch, i, kChF + q);
                    Fb[q / 3][q % 3] = adj.F[q * n + pid];
                }
                // F' = G F, G = I + dt C': dL/dF = G^T dL/dF', dL/dC' += dt dL/dF' F^T
                float Fn[3][3];
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        float s = 0.0f, u = 0.0f;
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            s += ((k == r ? 1.0f : 0.0f) + sc.dt * Cn[k][r]) * Fb[k][c];
                            u += Fb[r][k] * F[c][k];
                        }
                        Fn[r][c] = s;
                        Cb[r][c] += sc.dt * u;
                    }
#pragma unroll
                for (int q = 0; q < 9; ++q) adj.F[q * n + pid] = Fn[q / 3][q % 3];
            } else {
                // J' = (1 + dt tr C') J
                const float J = chan(A.ch, i, kChF), Jb = adj.F[pid];
                const float trc = Cn[0][0] + Cn[1][1] + Cn[2][2];
                adj.F[pid] = (1.0f + sc.dt * trc) * Jb;
#pragma unroll
                for (int k = 0; k < 3; ++k) Cb[k][k] += sc.dt * J * Jb;
            }
            // C' = D^-1 sum W v_i d^T: payload B = D^-1 dL/dC'; v' = sum W v_i: payload a = dL/dv'
            float Bq[3][3];
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) Bq[r][c] = sc.Dinv * Cb[r][c];
            // x through the weights: sum_i (v_i . (a + B d_i)) grad W_i; through d_i = x_i - x: -B^T v'
            float dw[3][3];
#pragma unroll
            for (int k = 0; k < 3; ++k) bspline_dw(ar.f[k], dw[k]);
            float gx[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll 1
            for (int ox = 0; ox < 3; ++ox)
#pragma unroll
                for (int oy = 0; oy < 3; ++oy)
#pragma unroll
                    for (int oz = 0; oz < 3; ++oz) {
                        const float4 v4 = s_v[((lc[0] + ox) * kTileSide + (lc[1] + oy)) * kTileSide + lc[2] + oz];
                        const float d[3] = {((float)ox - ar.f[0]) * sc.dx, ((float)oy - ar.f[1]) * sc.dx,
                                            ((float)oz - ar.f[2]) * sc.dx};
                        const float vi[3] = {v4.x, v4.y, v4.z};
                        float s = 0.0f;
#pragma unroll
                        for (int r = 0; r < 3; ++r)
                            s += vi[r] * (vb[r] + Bq[r][0] * d[0] + Bq[r][1] * d[1] + Bq[r][2] * d[2]);
                        const float wx = ar.w[0][ox], wy = ar.w[1][oy], wz = ar.w[2][oz];
                        gx[0] += s * dw[0][ox] * wy * wz;
                        gx[1] += s * wx * dw[1][oy] * wz;
                        gx[2] += s * wx * wy * dw[2][oz];
                    }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float bt = Bq[0][k] * vel[0] + Bq[1][k] * vel[1] + Bq[2][k] * vel[2];
                adj.x[k * n + pid] = xb[k] + sc.dxinv * gx[k] - bt;
                pay_a[k * stride + i] = vb[k];
            }
#pragma unroll
            for (int q = 0; q < 9; ++q) pay_B[q * stride + i] = Bq[q / 3][q % 3];
        }
        __syncthreads();
    }
}

// Grid adjoint, in place on gbar: channels 1-3 hold dL/dv of the updated
// velocities on entry and dL/dm, dL/dmv (channel 0, 1-3) on exit.  grid holds the
// combined mass and momentum.  Sticky-box nodes did not depend on mv.
__global__ void __launch_bounds__(kThreads) k_grid_adjoint(Partition P, Lattice L, const float *grid, float *gbar) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lx = lane >> 4, ly = (lane >> 2) & 3, lz = lane & 3;
    const uint32_t ng = P.counts[1];
    for (uint32_t gb = blockIdx.x * 4 + w; gb < ng; gb += gridDim.x * 4) {
        const int64_t l = P.gblk_lat[gb];
        const SceneParams sc = scene_params(L, L.nscenes == 1 ? 0u : P.gblk_scene[gb]);
        int Lx, Ly, Lz;
        lattice_coords(sc, l, Lx, Ly, Lz);
        const float *g = grid + (int64_t)gb * 256;
        float *o = gbar + (int64_t)gb * 256;
        const float m = g[lane], p0 = g[64 + lane], p1 = g[128 + lane], p2 = g[192 + lane];
        float b0 = o[64 + lane], b1 = o[128 + lane], b2 = o[192 + lane];
        if (sc.boundary == ZMPM_BOUNDARY_STICKY_BOX) {
            const int gx = (Lx + L.origin) * 4 + lx, gy = (Ly + L.origin) * 4 + ly, gz = (Lz + L.origin) * 4 + lz;
            const int lo = sc.pad, hi = sc.n_cells - sc.pad;
            if (gx < lo || gy < lo || gz < lo || gx > hi || gy > hi || gz > hi) b0 = b1 = b2 = 0.0f;
        }
        float mb = 0.0f;
        if (m != 0.0f) {
            // v = mv / m + dt g
            const float inv = 1.0f / m;
            mb = -(b0 * p0 + b1 * p1 + b2 * p2) * inv * inv;
            b0 *= inv;
            b1 *= inv;
            b2 *= inv;
        } else {
            b0 = b1 = b2 = 0.0f;
        }
        o[lane] = mb;
        o[64 + lane] = b0;
        o[128 + lane] = b1;
        o[192 + lane] = b2;
    }
}

// P2G adjoint.  In: the resident state, gbar = (dL/dm, dL/dmv) of the combined
// grid, adj.x and adj.F with the G2P part.  Out: adj.x and adj.F complete,
// adj.v and adj.C written, adj.mu / adj.lam accumulated (fixed corotated).
__global__ void __launch_bounds__(kAdjThreads) k_p2g_adjoint(State A, Partition P, Lattice L, const DevMaterial *mats,
                                                          const float *gbar, AdjArrays adj) {
    __shared__ float4 s_g[kTileNodes];
    const int t = threadIdx.x;
    const uint32_t nblocks = P.counts[0];
    const int64_t n = adj.n;
    const uint32_t chunk = blockIdx.x % kAdjChunks;
    for (uint32_t b = blockIdx.x / kAdjChunks; b < nblocks; b += gridDim.x / kAdjChunks) {
        const int64_t l = P.pblk_lat[b];
        const SceneParams sc = scene_params(L, L.nscenes == 1 ? 0u : P.pblk_scene[b]);
        const int64_t nb = sc.nb, nb2 = nb * nb;
        int lb[3];
        lattice_coords(sc, l, lb[0], lb[1], lb[2]);
        const int bc[3] = {(lb[0] + L.origin) * 4, (lb[1] + L.origin) * 4, (lb[2] + L.origin) * 4};
        stage_block_grid4(s_g, gbar, P.lat_gid, l, nb, nb2);
        __syncthreads();
        const uint32_t ps = P.pblk_start[b], pe = P.pblk_start[b + 1];
        for (uint32_t i = ps + chunk * kAdjThreads + t; i < pe; i += kAdjChunks * kAdjThreads) {
            float xp[3], vp[3], Cp[3][3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                xp[k] = chan(A.ch, i, kChX + k);
                vp[k] = chan(A.ch, i, kChV + k);
            }
#pragma unroll
            for (int q = 0; q < 9; ++q) Cp[q / 3][q % 3] = chan(A.ch, i, kChC + q);
            const uint32_t mw = chan_u(A.ch, i, kChMat), pid = chan_u(A.ch, i, kChPid);
            const DevMaterial mt = mats[mw & kMatIndexMask];
            const bool fcr = mt.model == ZMPM_MODEL_FCR;
            // the particle's own m, vol when it carries them (Transfer.cu:857-858)
            const float m = L.pp_mass ? chan(A.ch, i, kChMass) : mt.mass;
            const float ks = -sc.dt * sc.Dinv * (L.pp_mass ? chan(A.ch, i, kChVol) : mt.vol);  // sigma = ks P F^T (Transfer.cu:869)
            float F[3][3], Pk[3][3], cof[3][3], J = 0.0f, sig[3][3];
            double R[3][3];
            if (fcr) {
#pragma unroll
                for (int q = 0; q < 9; ++q) F[q / 3][q % 3] = chan(A.ch, i, kChF + q);
                fcr_piola_parts(F, L.pp_lame ? chan(A.ch, i, kChMu) : mt.mu, L.pp_lame ? chan(A.ch, i, kChLam) : mt.lam,
                                Pk, R, J, cof);
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        sig[r][c] = (ks * Pk[r][0]) * F[c][0] + (ks * Pk[r][1]) * F[c][1] + (ks * Pk[r][2]) * F[c][2];
            } else {
                J = chan(A.ch, i, kChF);
                const float tau = mt.bulk * J * (J - 1.0f);
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c) sig[r][c] = (r == c) ? ks * tau : 0.0f;
            }
            float Bm[3][3];
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) Bm[r][c] = fmaf(m, Cp[r][c], sig[r][c]);
            Arena ar;
            make_arena(xp, sc.dxinv, ar);
            const int lc[3] = {clamp03(ar.base[0] - bc[0]), clamp03(ar.base[1] - bc[1]), clamp03(ar.base[2] - bc[2])};
            float dw[3][3];
#pragma unroll
            for (int k = 0; k < 3; ++k) bspline_dw(ar.f[k], dw[k]);
            float G0[3] = {0.0f, 0.0f, 0.0f}, Bb[3][3] = {{0.0f}}, gx[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll 1
            for (int ox = 0; ox < 3; ++ox)
#pragma unroll
                for (int oy = 0; oy < 3; ++oy)
#pragma unroll
                    for (int oz = 0; oz < 3; ++oz) {
                        const float4 g4 = s_g[((lc[0] + ox) * kTileSide + (lc[1] + oy)) * kTileSide + lc[2] + oz];
                        const float d[3] = {((float)ox - ar.f[0]) * sc.dx, ((float)oy - ar.f[1]) * sc.dx,
                                            ((float)oz - ar.f[2]) * sc.dx};
                        const float wx = ar.w[0][ox], wy = ar.w[1][oy], wz = ar.w[2][oz];
                        const float W = wx * wy * wz;
                        const float mvb[3] = {g4.y, g4.z, g4.w};
                        // mass and momentum the particle sends to node i (for the weight gradient)
                        float q = g4.x * m;
#pragma unroll
                        for (int r = 0; r < 3; ++r) {
                            const float mom = m * vp[r] + Bm[r][0] * d[0] + Bm[r][1] * d[1] + Bm[r][2] * d[2];
                            q += mvb[r] * mom;
                            G0[r] += W * mvb[r];
#pragma unroll
                            for (int c = 0; c < 3; ++c) Bb[r][c] += W * mvb[r] * d[c];
                        }
                        gx[0] += q * dw[0][ox] * wy * wz;
                        gx[1] += q * wx * dw[1][oy] * wz;
                        gx[2] += q * wx * wy * dw[2][oz];
                    }
            // x: weight gradient, and d_i = x_i - x contributes -B^T G0
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float bt = Bm[0][k] * G0[0] + Bm[1][k] * G0[1] + Bm[2][k] * G0[2];
                adj.x[k * n + pid] += sc.dxinv * gx[k] - bt;
                adj.v[k * n + pid] = m * G0[k];
            }
#pragma unroll
            for (int q = 0; q < 9; ++q) adj.C[q * n + pid] = m * Bb[q / 3][q % 3];
            // sigma adjoint = Bb
            if (fcr) {
                // sigma = ks P F^T: dL/dP = ks Bb F, dL/dF += ks Bb^T P
                float Pb[3][3], Fb[3][3];
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        Pb[r][c] = ks * (Bb[r][0] * F[0][c] + Bb[r][1] * F[1][c] + Bb[r][2] * F[2][c]);
                        Fb[r][c] = ks * (Bb[0][r] * Pk[0][c] + Bb[1][r] * Pk[1][c] + Bb[2][r] * Pk[2][c]);
                    }
                float mub = 0.0f, lamb = 0.0f;
                fcr_piola_adjoint(F, R, J, cof, L.pp_lame ? chan(A.ch, i, kChMu) : mt.mu,
                                  L.pp_lame ? chan(A.ch, i, kChLam) : mt.lam, Pb, Fb, mub, lamb);
#pragma unroll
                for (int q = 0; q < 9; ++q) adj.F[q * n + pid] += Fb[q / 3][q % 3];
                adj.mu[pid] += mub;
                adj.lam[pid] += lamb;
            } else {
                // sigma = ks K J (J - 1) I
                adj.F[pid] += ks * mt.bulk * (2.0f * J - 1.0f) * (Bb[0][0] + Bb[1][1] + Bb[2][2]);
            }
        }
        __syncthreads();
    }
}

inline unsigned blocks_grid(int64_t max_blocks, int per_wg) {
    int64_t g = (max_blocks + per_wg - 1) / per_wg;
    if (g > (1 << 20)) g = 1 << 20;
    if (g < 1) g = 1;
    return (unsigned)g;
}

// kAdjChunks workgroups per particle block
inline unsigned chunked_grid(int64_t max_blocks) {
    int64_t g = max_blocks < 1 ? 1 : max_blocks;
    if (g > (1 << 17)) g = 1 << 17;
    return (unsigned)(g * kAdjChunks);
}

}  // namespace

void launch_g2p_adjoint(State A, Partition P, Lattice L, const float *gridv, AdjArrays adj, float *pay_a, float *pay_B,
                        int64_t stride, int64_t max_blocks, hipStream_t s) {
    if (max_blocks <= 0) return;
    hipLaunchKernelGGL(k_g2p_adjoint, dim3(chunked_grid(max_blocks)), dim3(kAdjThreads), 0, s, A, P, L, gridv, adj,
                       pay_a, pay_B, stride);
}

void launch_grid_adjoint(Partition P, Lattice L, const float *grid, float *gbar, int64_t max_blocks, hipStream_t s) {
    if (max_blocks <= 0) return;
    hipLaunchKernelGGL(k_grid_adjoint, dim3(blocks_grid(max_blocks, 4)), dim3(kThreads), 0, s, P, L, grid, gbar);
}

void launch_p2g_gather_adjoint(State A, Partition P, Lattice L, const DevMaterial *mats, const float *gbar,
                               AdjArrays adj, int64_t max_blocks, hipStream_t s) {
    if (max_blocks <= 0) return;
    hipLaunchKernelGGL(k_p2g_adjoint, dim3(chunked_grid(max_blocks)), dim3(kAdjThreads), 0, s, A, P, L, mats, gbar,
                       adj);
}

}  // namespace zmpm
