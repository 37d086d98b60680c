// CPU oracle for the MLS/APIC-MPM substep -- TEST INFRASTRUCTURE ONLY.
//
// This file is the checker, never the product: only tests/, the smoke check
// of __graft_entry__.py and the cpu_baseline leg of bench.py may load it.
// The product path (libzmpm.so) never links or calls it.
//
// It restates, in plain C++ templated on the scalar type (float = the FP32
// reference path, double = the FP64 oracle), the arithmetic of Zeno's
// CuLagrange MPM nodes:
//   partition   projects/CuLagrange/mpm/MPMPipeline.cu:75-84   (block key)
//   expand      projects/CuLagrange/mpm/MPMPipeline.cu:131-199 (halo blocks)
//   grid clear  projects/CuLagrange/mpm/MPMPipeline.cu:251-262
//   P2G (apic)  projects/CuLagrange/mpm/Transfer.cu:844-890
//   grid update projects/CuLagrange/mpm/MPMPipeline.cu:852-877
//   boundary    projects/CuLagrange/mpm/Boundary.cu:104-116 (sticky box)
//   G2P (apic)  projects/CuLagrange/mpm/Transfer.cu:1137-1173
//   model       projects/CuLagrange/mpm/Generation.cu:18-42 (E, nu -> FCR)
//   init        projects/CuLagrange/mpm/Generation.cpp:943-990
// The zpc primitives those call sites use (B-spline arena, FixedCorotated,
// SVD) are absent from the reference snapshot; they are restated from the
// published methods named at each function below.
//
// Determinism: P2G visits particle blocks in 8 colour classes (block coord
// parity); blocks of one colour have disjoint stencils, so they run in
// parallel without atomics and every grid node receives its contributions
// in a fixed order (block order, then particle index order).  Results are
// independent of the OpenMP thread count.
//
// Build: oracle/Makefile (g++ -O2 -fopenmp -ffp-contract=off).

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

enum { MODEL_FCR = 0, MODEL_JFLUID = 1, MODEL_STVK_HENCKY = 2, MODEL_NEO_HOOKEAN = 3, MODEL_STABLE_NEO_HOOKEAN = 4 };
enum { PLASTIC_NONE = 0, PLASTIC_DRUCKER_PRAGER = 1, PLASTIC_VON_MISES = 2, PLASTIC_CAM_CLAY = 3 };
enum { BOUNDARY_NONE = 0, BOUNDARY_STICKY_BOX = 1 };

struct Material {   // plain C layout shared with the Python binding
    int32_t model;  // MODEL_*
    float E, nu, rho, vol, bulk;
    int32_t plasticity;   // PLASTIC_* (MODEL_STVK_HENCKY)
    float friction_angle; // degrees
    float yield_stress;
    float flip_beta;      // ASFLIP positional adjustment (0: AFLIP): the particles' initial beta
    float nacc_beta;      // Cam-clay cohesion ratio ("beta", Generation.cu:80)
    float nacc_xi;        // Cam-clay hardening ("xi", Generation.cu:81)
    int32_t aniso;        // ANISO_* ("aniso", Generation.cu:60-66)
    float aniso_strength; // AnisotropicArap's strength (Generation.cu:62, default 10)
};
enum { ANISO_NONE = 0, ANISO_ARAP = 1 };
enum { TRANSFER_APIC = 0, TRANSFER_AFLIP = 1, TRANSFER_FLIP = 2 };
// ToZSParticles' initial logJp (Generation.cpp:987)
static const double kLogJpInit = -0.04;
// FLIP blend of the affine FLIP transfers (Transfer.cu:1199, "constexpr float flip = 0.99f")
static const double kFlipBlend = 0.99;

// A level-set collider, the C layout of zmpm_collider (the SDF samples on the host).
struct Collider {
    int32_t type, shape, scene;
    int32_t dims[3];
    float shape_params[6];
    const float *sdf;
    float origin[3];
    float spacing;
    float translation[3], translation_rate[3];
    float scale, scale_rate;
    float ypr_degrees[3];
    float angular_velocity[3];
    const int32_t *leaf_origins;  // SDF_LEAVES: (nleaves, 3) index coordinates of each leaf's first sample
    int32_t nleaves;
    int32_t reserved;
};
enum { COLLIDER_STICKY = 0, COLLIDER_SLIP = 1, COLLIDER_SEPARATE = 2 };
enum { SHAPE_HALFSPACE = 0, SHAPE_SPHERE = 1, SHAPE_BOX = 2, SHAPE_SDF_GRID = 3, SHAPE_SDF_LEAVES = 4 };

// The leaves of an SDF_LEAVES collider as a sorted (leaf coordinate -> leaf) table, searched by
// bisection (the library keeps a dense map over the leaves' bounding box instead).
struct LeafIndex {
    std::vector<std::pair<std::array<int32_t, 3>, int32_t>> leaves;
    explicit LeafIndex(const Collider &c) {
        if (c.shape != SHAPE_SDF_LEAVES) return;
        for (int32_t l = 0; l < c.nleaves; ++l) {
            const int32_t *o = c.leaf_origins + 3 * (int64_t)l;
            // origins are multiples of 8: an arithmetic shift is the floor division
            leaves.push_back({{o[0] >> 3, o[1] >> 3, o[2] >> 3}, l});
        }
        std::sort(leaves.begin(), leaves.end());
    }
    // leaf holding sample (i, j, k), or -1
    int32_t find(int64_t i, int64_t j, int64_t k) const {
        const std::array<int32_t, 3> key = {(int32_t)(i >> 3), (int32_t)(j >> 3), (int32_t)(k >> 3)};
        auto it = std::lower_bound(leaves.begin(), leaves.end(), std::make_pair(key, (int32_t)-1));
        return it != leaves.end() && it->first == key ? it->second : -1;
    }
};

struct Params {
    double dx, dt;
    double gravity[3];
    int32_t n_cells;  // domain side in cells (round(1/dx) for the unit box)
    int32_t pad;      // sticky wall thickness in cells
    int32_t boundary; // BOUNDARY_*
    int32_t threads;  // OpenMP threads (0 = runtime default)
};

// ---------------------------------------------------------------- 3x3 helpers
template <typename T> struct M3 {
    T a[3][3];
};

// Symmetric eigen-decomposition of a 3x3 matrix by the cyclic Jacobi method
// (Golub & Van Loan, "Matrix Computations", 4th ed., sec. 8.5.2-8.5.3:
// sym.schur2 rotations applied cyclically over (0,1),(0,2),(1,2)).
// On return A is (numerically) diagonal and A_in = V diag(A) V^T.
static void jacobi_eig3(double A[3][3], double V[3][3]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) V[i][j] = (i == j) ? 1.0 : 0.0;
    static const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
    for (int sweep = 0; sweep < 16; ++sweep) {
        double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        double dia = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
        if (off <= 1e-34 * dia || off == 0.0) break;
        for (int r = 0; r < 3; ++r) {
            const int p = P[r], q = Q[r];
            const double apq = A[p][q];
            if (apq == 0.0) continue;
            const double tau = (A[q][q] - A[p][p]) / (2.0 * apq);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
            const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
            A[p][p] -= t * apq;
            A[q][q] += t * apq;
            A[p][q] = A[q][p] = 0.0;
            const int k = 3 - p - q;  // the remaining index
            const double akp = A[k][p], akq = A[k][q];
            A[k][p] = A[p][k] = c * akp - s * akq;
            A[k][q] = A[q][k] = s * akp + c * akq;
            for (int i = 0; i < 3; ++i) {
                const double vp = V[i][p], vq = V[i][q];
                V[i][p] = c * vp - s * vq;
                V[i][q] = s * vp + c * vq;
            }
        }
    }
}

// Signed SVD F = U diag(s) V^T with U, V proper rotations, s0 >= s1 >= |s2|
// and sign(s2) = sign(det F): eigen-decomposition of F^T F for V, then a
// Givens QR of F V for U (McAdams, Selle, Tamstorf, Teran, Sifakis 2011,
// "Computing the Singular Value Decomposition of 3x3 matrices with minimal
// branching and elementary floating point operations", UW-Madison TR1690,
// sec. 2-4, here with exact rotations in FP64).  The inversion convention
// (negate the smallest singular value) is that of Irving, Teran, Fedkiw 2004.
static void svd3(const double F[3][3], double U[3][3], double s[3], double V[3][3]) {
    double S[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            S[i][j] = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
    jacobi_eig3(S, V);
    // sort eigenvalues (and V columns) descending
    double ev[3] = {S[0][0], S[1][1], S[2][2]};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2 - i; ++j)
            if (ev[j] < ev[j + 1]) {
                std::swap(ev[j], ev[j + 1]);
                for (int r = 0; r < 3; ++r) std::swap(V[r][j], V[r][j + 1]);
            }
    const double detV = V[0][0] * (V[1][1] * V[2][2] - V[1][2] * V[2][1]) -
                        V[0][1] * (V[1][0] * V[2][2] - V[1][2] * V[2][0]) +
                        V[0][2] * (V[1][0] * V[2][1] - V[1][1] * V[2][0]);
    if (detV < 0.0)
        for (int r = 0; r < 3; ++r) V[r][2] = -V[r][2];
    double B[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) B[i][j] = F[i][0] * V[0][j] + F[i][1] * V[1][j] + F[i][2] * V[2][j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) U[i][j] = (i == j) ? 1.0 : 0.0;
    // Givens rotations zeroing B[1][0], B[2][0], B[2][1]; B <- G B, U <- U G^T
    static const int GP[3] = {0, 0, 1}, GQ[3] = {1, 2, 2}, GC[3] = {0, 0, 1};
    for (int g = 0; g < 3; ++g) {
        const int p = GP[g], q = GQ[g], col = GC[g];
        const double a = B[p][col], b = B[q][col];
        const double r = std::sqrt(a * a + b * b);
        double c = 1.0, sn = 0.0;
        if (r > 0.0) { c = a / r; sn = b / r; }
        for (int j = 0; j < 3; ++j) {
            const double bp = B[p][j], bq = B[q][j];
            B[p][j] = c * bp + sn * bq;
            B[q][j] = -sn * bp + c * bq;
        }
        for (int i = 0; i < 3; ++i) {
            const double up = U[i][p], uq = U[i][q];
            U[i][p] = c * up + sn * uq;
            U[i][q] = -sn * up + c * uq;
        }
    }
    s[0] = B[0][0]; s[1] = B[1][1]; s[2] = B[2][2];
}

// Rotation factor R of F = R S (R proper): R = U V^T from the signed SVD.
static void polar_rotation(const double F[3][3], double R[3][3]) {
    double U[3][3], s[3], V[3][3];
    svd3(F, U, s, V);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) R[i][j] = U[i][0] * V[j][0] + U[i][1] * V[j][1] + U[i][2] * V[j][2];
}

// Lame parameters from Young's modulus and Poisson ratio (the conversion
// FixedCorotated<float>{E, nu} performs, Generation.cu:41).
template <typename T> static void lame(T E, T nu, T &mu, T &lam) {
    mu = E / (T(2) * (T(1) + nu));
    lam = E * nu / ((T(1) + nu) * (T(1) - T(2) * nu));
}

template <typename T> static T det3(const T F[3][3]) {
    return F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) - F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0]) +
           F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]);
}

// Fixed-corotated first Piola-Kirchhoff stress (Stomakhin, Howes, Schroeder,
// Teran 2012): P = 2 mu (F - R) + lam (J - 1) cof(F), cof(F) = J F^{-T}.
// The rotation R comes from the FP64 signed SVD; the rest is in T.
template <typename T> static void fcr_piola(const T F[3][3], T mu, T lam, T P[3][3]) {
    double Fd[3][3], Rd[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Fd[i][j] = (double)F[i][j];
    polar_rotation(Fd, Rd);
    const T J = det3(F);
    T cof[3][3];
    cof[0][0] = F[1][1] * F[2][2] - F[1][2] * F[2][1];
    cof[0][1] = F[1][2] * F[2][0] - F[1][0] * F[2][2];
    cof[0][2] = F[1][0] * F[2][1] - F[1][1] * F[2][0];
    cof[1][0] = F[0][2] * F[2][1] - F[0][1] * F[2][2];
    cof[1][1] = F[0][0] * F[2][2] - F[0][2] * F[2][0];
    cof[1][2] = F[0][1] * F[2][0] - F[0][0] * F[2][1];
    cof[2][0] = F[0][1] * F[1][2] - F[0][2] * F[1][1];
    cof[2][1] = F[0][2] * F[1][0] - F[0][0] * F[1][2];
    cof[2][2] = F[0][0] * F[1][1] - F[0][1] * F[1][0];
    const T twomu = T(2) * mu, lj = lam * (J - T(1));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) P[i][j] = twomu * (F[i][j] - (T)Rd[i][j]) + lj * cof[i][j];
}

// Neo-Hookean ("nhk", the MPM course of Jiang et al. 2016, eq. 48):
// psi = mu/2 (|F|^2 - 3) - mu log J + lam/2 (log J)^2, P = mu (F - F^-T) + lam log(J) F^-T,
// with J clamped to >= 1e-6; stable Neo-Hookean ("snhk", Smith, de Goes, Kim 2018, without the
// log term): psi = mu/2 (|F|^2 - 3) - mu (J - 1) + lam/2 (J - 1)^2, P = mu F + (lam (J - 1) - mu) cof F.
template <typename T> static void neo_hookean_piola(const T F[3][3], T mu, T lam, bool stable, T P[3][3]) {
    T cof[3][3];
    cof[0][0] = F[1][1] * F[2][2] - F[1][2] * F[2][1];
    cof[0][1] = F[1][2] * F[2][0] - F[1][0] * F[2][2];
    cof[0][2] = F[1][0] * F[2][1] - F[1][1] * F[2][0];
    cof[1][0] = F[0][2] * F[2][1] - F[0][1] * F[2][2];
    cof[1][1] = F[0][0] * F[2][2] - F[0][2] * F[2][0];
    cof[1][2] = F[0][1] * F[2][0] - F[0][0] * F[2][1];
    cof[2][0] = F[0][1] * F[1][2] - F[0][2] * F[1][1];
    cof[2][1] = F[0][2] * F[1][0] - F[0][0] * F[1][2];
    cof[2][2] = F[0][0] * F[1][1] - F[0][1] * F[1][0];
    const T J = F[0][0] * cof[0][0] + F[0][1] * cof[0][1] + F[0][2] * cof[0][2];
    T k;
    if (stable) {
        k = lam * (J - T(1)) - mu;
    } else {
        const T Jc = std::max(J, T(1e-6));
        k = (lam * std::log(Jc) - mu) / Jc;
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) P[i][j] = mu * F[i][j] + k * cof[i][j];
}

// Anisotropic ARAP fiber term ("arap": AnisotropicArap{E, nu, strength},
// Generation.cu:60-63, added to P at Transfer.cu:866-867; zpc's
// AnisotropicArap is absent).  The energy of Kim, De Goes, Iben 2019,
// "Anisotropic Elasticity for Inversion-Safety and Element Rehabilitation":
// psi = mu_a/2 (sqrt(I5) - S(I4))^2, I5 = |F a|^2, I4 = a^T S a with S = R^T F
// (R the proper polar rotation), S(I4) = -1 if I4 < 0 else 1, mu_a = strength mu;
// P = mu_a (1 - S(I4) / sqrt(I5)) F a a^T with sqrt(I5) clamped to >= 1e-6.
// a is the particle's rest-space fiber, used as given.  S is positive definite
// when det F > 0, so only det F <= 0 needs the rotation.
template <typename T> static void arap_fiber_add(const T F[3][3], const T a[3], T mu_a, T P[3][3]) {
    T Fa[3];
    for (int r = 0; r < 3; ++r) Fa[r] = F[r][0] * a[0] + F[r][1] * a[1] + F[r][2] * a[2];
    double Fd[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Fd[i][j] = (double)F[i][j];
    T sgn = T(1);
    if (!(det3(Fd) > 0.0)) {
        double Rd[3][3], I4 = 0.0;
        polar_rotation(Fd, Rd);
        for (int r = 0; r < 3; ++r)
            I4 += (Rd[r][0] * (double)a[0] + Rd[r][1] * (double)a[1] + Rd[r][2] * (double)a[2]) * (double)Fa[r];
        if (I4 < 0.0) sgn = T(-1);
    }
    const T r5 = std::max(std::sqrt(Fa[0] * Fa[0] + Fa[1] * Fa[1] + Fa[2] * Fa[2]), T(1e-6));
    const T k = mu_a * (T(1) - sgn / r5);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) P[r][c] += (k * Fa[r]) * a[c];
}

// StVK with Hencky strain (Klar, Gast, Pradhana, Fu, Schroeder, Jiang, Teran
// 2016, sec. 4), psi = mu |e|^2 + lam/2 tr(e)^2 with e = log|s| from the signed
// SVD F = U diag(s) V^T (|s| clamped at 128 FLT_EPSILON, the clamp of the
// reference's curve strain, MPMPipeline.cu:1037).  Kirchhoff stress
// tau = P F^T = U diag(2 mu e + lam tr(e)) U^T, in FP64.
static const double kMinStretch = 128.0 * 1.1920928955078125e-7;

static void hencky(const double F[3][3], double U[3][3], double s[3], double V[3][3], double e[3]) {
    svd3(F, U, s, V);
    for (int k = 0; k < 3; ++k) e[k] = std::log(std::max(std::fabs(s[k]), kMinStretch));
}

static void stvk_hencky_tau(const double F[3][3], double mu, double lam, double tau[3][3]) {
    double U[3][3], s[3], V[3][3], e[3];
    hencky(F, U, s, V, e);
    const double tr = e[0] + e[1] + e[2];
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = 2.0 * mu * e[k] + lam * tr;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) tau[i][j] = U[i][0] * d[0] * U[j][0] + U[i][1] * d[1] * U[j][1] + U[i][2] * d[2] * U[j][2];
}

// Drucker-Prager friction coefficient alpha = sqrt(2/3) 2 sin(phi) / (3 - sin(phi))
// (Klar et al. 2016, eq. 27, no cohesion), phi in degrees.
static double dp_alpha(double phi_deg) {
    const double sp = std::sin(phi_deg * 3.14159265358979323846 / 180.0);
    return std::sqrt(2.0 / 3.0) * 2.0 * sp / (3.0 - sp);
}

// Return mapping of the Hencky strain (ZSReturnMapping, MPMPipeline.cu:953-972).
// e^ = e - tr(e)/3.  Drucker-Prager (Klar et al. 2016, sec. 7.3, non-associative):
// tr(e) >= 0 -> e' = 0; else dg = |e^| + (3 lam + 2 mu)/(2 mu) tr(e) alpha,
// e' = e - dg e^/|e^| when dg > 0.  Von Mises (Jiang et al. 2016 MPM course,
// sec. 9): dg = |e^| - yield/(2 mu), same e'.  F' = U diag(sign(s) exp(e')) V^T.
// Returns 1 if F changed.
static int project_strain(double F[3][3], double mu, double lam, int plasticity, double alpha, double ys) {
    if (plasticity == PLASTIC_NONE) return 0;
    double U[3][3], s[3], V[3][3], e[3];
    hencky(F, U, s, V, e);
    const double tr = e[0] + e[1] + e[2];
    double eh[3], nh = 0.0;
    for (int k = 0; k < 3; ++k) {
        eh[k] = e[k] - tr / 3.0;
        nh += eh[k] * eh[k];
    }
    nh = std::sqrt(nh);
    double ep[3];
    if (plasticity == PLASTIC_DRUCKER_PRAGER) {
        if (tr >= 0.0) {
            ep[0] = ep[1] = ep[2] = 0.0;
        } else {
            const double dg = nh + (3.0 * lam + 2.0 * mu) / (2.0 * mu) * tr * alpha;
            if (!(dg > 0.0)) return 0;
            for (int k = 0; k < 3; ++k) ep[k] = e[k] - dg / nh * eh[k];
        }
    } else {
        const double dg = nh - ys / (2.0 * mu);
        if (!(dg > 0.0)) return 0;
        for (int k = 0; k < 3; ++k) ep[k] = e[k] - dg / nh * eh[k];
    }
    double sp[3];
    for (int k = 0; k < 3; ++k) sp[k] = std::copysign(std::exp(ep[k]), s[k]);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) F[i][j] = U[i][0] * sp[0] * V[j][0] + U[i][1] * sp[1] * V[j][1] + U[i][2] * sp[2] * V[j][2];
    return 1;
}

// Non-associative Cam-clay ("nacc", Generation.cu:77-82; zpc's model is absent),
// the published model of Gaume, Gast, Teran, van Herwijnen, Jiang 2018
// ("Dynamic anticrack propagation in snow", Nat. Commun. 9:3047, Methods) on the
// StVK-Hencky Kirchhoff stress tau = U diag(2 mu e + lam tr(e)) U^T:
//   p = -tr(tau)/3, q = sqrt(3/2) |dev tau|, K = lam + 2 mu / 3,
//   M = 6 sin(phi) / (3 - sin(phi)), p0 = K (1e-5 + sinh(xi max(-logJp, 0))),
//   yield: (1 + 2 beta) q^2 + M^2 (p + beta p0)(p - p0) <= 0.
// Outside: p > p0 -> the compression tip (tau = -p0 I), p < -beta p0 -> the
// tension tip (tau = beta p0 I), else p kept and dev tau scaled onto the
// ellipse.  logJp += log(J_trial / J_new).  F' = U diag(sign(s) exp(e')) V^T.
// Returns 1 if F changed.  Written in terms of stresses (the GPU kernel works in
// strains); the two agree because tau is linear in e.
static int project_cam_clay(double F[3][3], double mu, double lam, double phi_deg, double beta, double xi,
                            double &logJp) {
    double U[3][3], s[3], V[3][3], e[3];
    hencky(F, U, s, V, e);
    const double K = lam + 2.0 * mu / 3.0;
    const double sp = std::sin(phi_deg * 3.14159265358979323846 / 180.0);
    const double M = 6.0 * sp / (3.0 - sp);
    const double tr = e[0] + e[1] + e[2];
    double tau[3], p = 0.0;
    for (int k = 0; k < 3; ++k) {
        tau[k] = 2.0 * mu * e[k] + lam * tr;
        p -= tau[k] / 3.0;
    }
    double dev[3], dn2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        dev[k] = tau[k] + p;
        dn2 += dev[k] * dev[k];
    }
    const double q2 = 1.5 * dn2;
    const double p0 = K * (1e-5 + std::sinh(xi * std::max(-logJp, 0.0)));
    const double y = (1.0 + 2.0 * beta) * q2 + M * M * (p + beta * p0) * (p - p0);
    if (!(y > 0.0)) return 0;
    double tau_n[3];
    if (p > p0) {
        for (int k = 0; k < 3; ++k) tau_n[k] = -p0;
    } else if (p < -beta * p0) {
        for (int k = 0; k < 3; ++k) tau_n[k] = beta * p0;
    } else {
        const double qn2 = std::max(-M * M * (p + beta * p0) * (p - p0) / (1.0 + 2.0 * beta), 0.0);
        const double r = std::sqrt(qn2 / q2);
        for (int k = 0; k < 3; ++k) tau_n[k] = -p + r * dev[k];
    }
    // invert tau = 2 mu e + lam tr(e): tr(e) = tr(tau) / (2 mu + 3 lam), e = (tau - lam tr(e)) / (2 mu)
    const double trt = tau_n[0] + tau_n[1] + tau_n[2];
    const double tre = trt / (2.0 * mu + 3.0 * lam);
    double en[3];
    for (int k = 0; k < 3; ++k) en[k] = (tau_n[k] - lam * tre) / (2.0 * mu);
    logJp += tr - (en[0] + en[1] + en[2]);
    double sn[3];
    for (int k = 0; k < 3; ++k) sn[k] = std::copysign(std::exp(en[k]), s[k]);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) F[i][j] = U[i][0] * sn[0] * V[j][0] + U[i][1] * sn[1] * V[j][1] + U[i][2] * sn[2] * V[j][2];
    return 1;
}

// ---------------------------------------------------------------- colliders
// Collider rotation R = Ry(yaw) Rx(pitch) Rz(roll) (degrees), frame -> world.
static void collider_rotation(const Collider &c, double R[3][3]) {
    const double r = 3.14159265358979323846 / 180.0;
    const double a = c.ypr_degrees[0] * r, b = c.ypr_degrees[1] * r, g = c.ypr_degrees[2] * r;
    const double Ry[3][3] = {{std::cos(a), 0, std::sin(a)}, {0, 1, 0}, {-std::sin(a), 0, std::cos(a)}};
    const double Rx[3][3] = {{1, 0, 0}, {0, std::cos(b), -std::sin(b)}, {0, std::sin(b), std::cos(b)}};
    const double Rz[3][3] = {{std::cos(g), -std::sin(g), 0}, {std::sin(g), std::cos(g), 0}, {0, 0, 1}};
    double T[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            T[i][j] = 0;
            for (int k = 0; k < 3; ++k) T[i][j] += Ry[i][k] * Rx[k][j];
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            R[i][j] = 0;
            for (int k = 0; k < 3; ++k) R[i][j] += T[i][k] * Rz[k][j];
        }
}

// Signed distance and gradient of the collider's shape in its own frame.
static double shape_phi(const Collider &c, const LeafIndex &li, const double X[3], double g[3]) {
    const float *p = c.shape_params;
    switch (c.shape) {
    case SHAPE_HALFSPACE: {
        const double nn = std::sqrt((double)p[0] * p[0] + (double)p[1] * p[1] + (double)p[2] * p[2]);
        for (int k = 0; k < 3; ++k) g[k] = p[k] / nn;
        return (p[0] * X[0] + p[1] * X[1] + p[2] * X[2] - p[3]) / nn;
    }
    case SHAPE_SPHERE: {
        const double d[3] = {X[0] - p[0], X[1] - p[1], X[2] - p[2]};
        const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (int k = 0; k < 3; ++k) g[k] = r > 0 ? d[k] / r : (k == 1 ? 1.0 : 0.0);
        return r - p[3];
    }
    case SHAPE_BOX: {
        // distance to an axis-aligned box: outside the Euclidean distance to its
        // nearest point, inside minus the distance to its nearest face
        double out = 0, inside = -1e300;
        int face = 0;
        double q[3], sg[3];
        for (int k = 0; k < 3; ++k) {
            const double cen = 0.5 * ((double)p[k] + p[3 + k]), half = 0.5 * ((double)p[3 + k] - p[k]);
            sg[k] = X[k] < cen ? -1.0 : 1.0;
            q[k] = std::fabs(X[k] - cen) - half;
            if (q[k] > 0) out += q[k] * q[k];
            if (q[k] > inside) { inside = q[k]; face = k; }
        }
        if (inside > 0) {
            out = std::sqrt(out);
            for (int k = 0; k < 3; ++k) g[k] = q[k] > 0 ? sg[k] * q[k] / out : 0.0;
            return out;
        }
        for (int k = 0; k < 3; ++k) g[k] = k == face ? sg[k] : 0.0;
        return inside;
    }
    case SHAPE_SDF_LEAVES: {
        // 8^3-sample leaves (OpenVDB's leaf nodes: sample (x, y, z) of the leaf at offset
        // (x & 7) * 64 + (y & 7) * 8 + (z & 7)); a sample outside every leaf reads the
        // background shape_params[0].  Trilinear in the samples at origin + spacing (i, j, k).
        int64_t i0[3];
        double f[3];
        for (int k = 0; k < 3; ++k) {
            const double u = (X[k] - c.origin[k]) / c.spacing;
            if (!(std::fabs(u) < 1e9)) {
                g[0] = g[1] = g[2] = 0;
                return p[0];
            }
            i0[k] = (int64_t)std::floor(u);
            f[k] = u - (double)i0[k];
        }
        auto at = [&](int a, int b, int e) {
            const int64_t i = i0[0] + a, j = i0[1] + b, k = i0[2] + e;
            const int32_t l = li.find(i, j, k);
            return l < 0 ? (double)p[0] : (double)c.sdf[(int64_t)l * 512 + (i & 7) * 64 + (j & 7) * 8 + (k & 7)];
        };
        double phi = 0;
        g[0] = g[1] = g[2] = 0;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int e = 0; e < 2; ++e) {
                    const double w[3] = {a ? f[0] : 1 - f[0], b ? f[1] : 1 - f[1], e ? f[2] : 1 - f[2]};
                    const double s = at(a, b, e);
                    phi += w[0] * w[1] * w[2] * s;
                    g[0] += (a ? 1 : -1) * w[1] * w[2] * s / c.spacing;
                    g[1] += (b ? 1 : -1) * w[0] * w[2] * s / c.spacing;
                    g[2] += (e ? 1 : -1) * w[0] * w[1] * s / c.spacing;
                }
        return phi;
    }
    default: {  // sampled: trilinear interpolation, gradient of the interpolant
        int i0[3];
        double f[3];
        for (int k = 0; k < 3; ++k) {
            const double u = (X[k] - c.origin[k]) / c.spacing;
            if (!(u >= 0 && u <= c.dims[k] - 1)) {
                g[0] = g[1] = g[2] = 0;
                return std::numeric_limits<double>::infinity();
            }
            i0[k] = std::min((int)u, c.dims[k] - 2);
            f[k] = u - i0[k];
        }
        auto at = [&](int a, int b, int e) {
            return (double)c.sdf[((int64_t)(i0[0] + a) * c.dims[1] + (i0[1] + b)) * c.dims[2] + (i0[2] + e)];
        };
        double phi = 0;
        g[0] = g[1] = g[2] = 0;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int e = 0; e < 2; ++e) {
                    const double w[3] = {a ? f[0] : 1 - f[0], b ? f[1] : 1 - f[1], e ? f[2] : 1 - f[2]};
                    const double s = at(a, b, e);
                    phi += w[0] * w[1] * w[2] * s;
                    g[0] += (a ? 1 : -1) * w[1] * w[2] * s / c.spacing;
                    g[1] += (b ? 1 : -1) * w[0] * w[2] * s / c.spacing;
                    g[2] += (e ? 1 : -1) * w[0] * w[1] * s / c.spacing;
                }
        return phi;
    }
    }
}

// The response of ApplyGridBoundaryOnZSGrid (Boundary.cu:55-66) for a node at x
// inside the collider (phi < 0), object velocity from the translation, rotation
// and scale rates.  Returns 1 if the node was inside.
template <typename T>
static int resolve_collision(const Collider &c, const LeafIndex &li, const double x[3], T v[3]) {
    double R[3][3];
    collider_rotation(c, R);
    double d[3], X[3], g[3];
    for (int k = 0; k < 3; ++k) d[k] = x[k] - c.translation[k];
    for (int k = 0; k < 3; ++k) X[k] = (R[0][k] * d[0] + R[1][k] * d[1] + R[2][k] * d[2]) / c.scale;
    const double phi = shape_phi(c, li, X, g) * c.scale;
    if (!(phi < 0)) return 0;
    const double om[3] = {c.angular_velocity[0], c.angular_velocity[1], c.angular_velocity[2]};
    double vo[3];
    const double cross[3] = {om[1] * d[2] - om[2] * d[1], om[2] * d[0] - om[0] * d[2], om[0] * d[1] - om[1] * d[0]};
    for (int k = 0; k < 3; ++k) vo[k] = c.translation_rate[k] + cross[k] + c.scale_rate / c.scale * d[k];
    if (c.type == COLLIDER_STICKY) {
        for (int k = 0; k < 3; ++k) v[k] = (T)vo[k];
        return 1;
    }
    double n[3], nn = 0;
    for (int k = 0; k < 3; ++k) {
        n[k] = R[k][0] * g[0] + R[k][1] * g[1] + R[k][2] * g[2];
        nn += n[k] * n[k];
    }
    nn = std::sqrt(nn);
    if (!(nn > 0)) return 1;
    double vr[3], proj = 0;
    for (int k = 0; k < 3; ++k) {
        n[k] /= nn;
        vr[k] = (double)v[k] - vo[k];
        proj += n[k] * vr[k];
    }
    if (c.type == COLLIDER_SLIP || proj < 0)
        for (int k = 0; k < 3; ++k) vr[k] -= proj * n[k];
    for (int k = 0; k < 3; ++k) v[k] = (T)(vr[k] + vo[k]);
    return 1;
}

// Quadratic B-spline arena of one particle (zpc make_local_arena, quadratic,
// collocated): X = x/dx, base = floor(X - 0.5), f = X - base in [0.5, 1.5),
// w0 = (1.5-f)^2/2, w1 = 0.75-(f-1)^2, w2 = (f-0.5)^2/2, diff = (o - f) dx.
template <typename T> struct Arena {
    int base[3];
    T w[3][3];
    T d[3][3];   // d[axis][o] = (o - f) * dx
    T dw[3][3];  // dw[axis][o] = d w / d f: f - 3/2, 2 (1 - f), f - 1/2 (the weight gradient times dx)
};

template <typename T> static inline int block_floor(T X) {
    // lower_trunc(x * dxinv - 0.5) with a double literal 0.5 (MPMPipeline.cu:77-83)
    return (int)std::floor((double)X - 0.5);
}

template <typename T> static inline void make_arena(const T xp[3], T dx, T dxinv, Arena<T> &a) {
    for (int k = 0; k < 3; ++k) {
        const T X = xp[k] * dxinv;
        a.base[k] = block_floor(X);
        const T f = X - (T)a.base[k];
        a.w[k][0] = T(0.5) * (T(1.5) - f) * (T(1.5) - f);
        a.w[k][1] = T(0.75) - (f - T(1)) * (f - T(1));
        a.w[k][2] = T(0.5) * (f - T(0.5)) * (f - T(0.5));
        for (int o = 0; o < 3; ++o) a.d[k][o] = ((T)o - f) * dx;
        a.dw[k][0] = f - T(1.5);
        a.dw[k][1] = T(2) * (T(1) - f);
        a.dw[k][2] = f - T(0.5);
    }
}

static inline int64_t pack_block(int bx, int by, int bz) {
    // bx, by, bz are block coordinates in units of blocks; lexicographic (x,y,z)
    const int64_t off = 1 << 20;
    return ((int64_t)(bx + off) << 42) | ((int64_t)(by + off) << 21) | (int64_t)(bz + off);
}
static inline void unpack_block(int64_t k, int &bx, int &by, int &bz) {
    const int64_t off = 1 << 20, m = (1 << 21) - 1;
    bx = (int)((k >> 42) & m) - (int)off;
    by = (int)((k >> 21) & m) - (int)off;
    bz = (int)(k & m) - (int)off;
}
static inline int fdiv4(int c) { return (c - (c & 3)) / 4; }  // floor(c / 4), two's complement

// ---------------------------------------------------------------- session
template <typename T> struct Session {
    Params prm;
    std::vector<Material> mats;
    int64_t n = 0;
    std::vector<T> x, v, C, F;  // SoA planes: x[k*n+i], C[(3r+c)*n+i]
    std::vector<int32_t> mat;
    // partition: particle blocks (sorted keys), particle order grouped by block
    std::vector<int64_t> pkeys;
    std::vector<int64_t> porder;  // particle indices, stable-sorted by block key
    std::vector<int64_t> pstart;  // pkeys.size()+1 offsets into porder
    // grid: blocks B + {0,1}^3, channel-major blocks (m, mv0, mv1, mv2) x 64 cells
    std::vector<int64_t> gkeys;
    std::unordered_map<int64_t, int32_t> gmap;
    std::vector<T> grid;
    T maxVelSqr = 0;
    int32_t status = 0;
    std::vector<Collider> colliders;
    std::vector<LeafIndex> leaf_index;  // one per collider (empty unless SDF_LEAVES)
    void set_colliders(const Collider *c, int32_t n) {
        colliders.assign(c, c + n);
        leaf_index.clear();
        for (int32_t q = 0; q < n; ++q) leaf_index.emplace_back(c[q]);
    }
    // AFLIP/ASFLIP (ZenoGrid transferScheme "aflip"/"asflip", Transfer.cu:935-984, 1174-1214):
    // gstar holds the "vstar" channel (stress momentum, then the updated velocity) beside grid's
    // (m, v), same layout (channel 0 unused)
    int32_t transfer = TRANSFER_APIC;
    std::vector<T> gstar;
    // boundary particles (ZSBoundaryPrimitiveToZSGrid + ApplyGridBoundaryOnZSGrid), SoA
    std::vector<double> bnd_x, bnd_v, bnd_m, bnd_nrm;
    int64_t bnd_n = 0;
    int32_t bnd_type = COLLIDER_STICKY;
    std::vector<T> pp_mu, pp_lam;  // per-particle Lame parameters (empty: the materials'), Transfer.cu:861-864
    std::vector<float> pp_m, pp_vol;  // per-particle mass and volume (empty: the materials'), Transfer.cu:857-858
    std::vector<T> beta;    // per particle: ASFLIP positional adjustment ("beta" channel)
    std::vector<T> logJp;   // per particle: Cam-clay plastic volume log ("logJp" channel)
    std::vector<T> fiber;   // SoA 3 x n: the rest-space fiber "a" (empty: (0, 1, 0), Generation.cpp:982)
    std::array<T, 3> fiber_of(int64_t i) const {
        if (fiber.empty()) return {T(0), T(1), T(0)};
        return {fiber[i], fiber[n + i], fiber[2 * n + i]};
    }
    bool flip_style() const { return transfer != TRANSFER_APIC; }

    int threads() const { return prm.threads > 0 ? prm.threads : 1; }

    void partition() {
        const T dxinv = T(1) / (T)prm.dx;
        std::vector<int64_t> key(n);
#pragma omp parallel for num_threads(threads()) schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            int c[3];
            for (int k = 0; k < 3; ++k) c[k] = block_floor(x[k * n + i] * dxinv);
            key[i] = pack_block(fdiv4(c[0]), fdiv4(c[1]), fdiv4(c[2]));
        }
        porder.resize(n);
        std::iota(porder.begin(), porder.end(), 0);
        std::stable_sort(porder.begin(), porder.end(), [&](int64_t a, int64_t b) { return key[a] < key[b]; });
        pkeys.clear();
        pstart.clear();
        for (int64_t j = 0; j < n; ++j) {
            const int64_t k = key[porder[j]];
            if (j == 0 || k != pkeys.back()) {
                pkeys.push_back(k);
                pstart.push_back(j);
            }
        }
        pstart.push_back(n);
        // expansion to B + {0,1}^3: the blocks a quadratic stencil can touch
        gkeys.clear();
        gkeys.reserve(pkeys.size() * 2);
        for (int64_t k : pkeys) {
            int bx, by, bz;
            unpack_block(k, bx, by, bz);
            for (int e = 0; e < 8; ++e) gkeys.push_back(pack_block(bx + (e >> 2), by + ((e >> 1) & 1), bz + (e & 1)));
        }
        std::sort(gkeys.begin(), gkeys.end());
        gkeys.erase(std::unique(gkeys.begin(), gkeys.end()), gkeys.end());
        gmap.clear();
        gmap.reserve(gkeys.size() * 2);
        for (size_t g = 0; g < gkeys.size(); ++g) gmap[gkeys[g]] = (int32_t)g;
        grid.assign(gkeys.size() * 4 * 64, T(0));
        gstar.assign(flip_style() ? gkeys.size() * 4 * 64 : 0, T(0));
    }

    // The 8 grid blocks a particle block's stencils can reach: e = (ex,ey,ez) in {0,1}^3.
    void neighbour_blocks(int64_t pkey, int32_t nb[8]) const {
        int bx, by, bz;
        unpack_block(pkey, bx, by, bz);
        for (int e = 0; e < 8; ++e) {
            auto it = gmap.find(pack_block(bx + (e >> 2), by + ((e >> 1) & 1), bz + (e & 1)));
            nb[e] = it == gmap.end() ? -1 : it->second;
        }
    }

    // grid node address for global node coordinate g (within the 2x2x2 block neighbourhood of pkey)
    inline T *node(const int32_t nb[8], int bx0, int by0, int bz0, int gx, int gy, int gz) {
        const int ex = fdiv4(gx) - bx0, ey = fdiv4(gy) - by0, ez = fdiv4(gz) - bz0;
        const int32_t g = nb[(ex << 2) | (ey << 1) | ez];
        const int cell = ((gx & 3) * 4 + (gy & 3)) * 4 + (gz & 3);
        return &grid[(size_t)g * 256 + cell];  // channel c at +64*c
    }

    void p2g() {
        const T dx = (T)prm.dx, dxinv = T(1) / dx, dt = (T)prm.dt;
        const T Dinv = T(4) * dxinv * dxinv;
        const int64_t nb_ = (int64_t)pkeys.size();
        for (int colour = 0; colour < 8; ++colour) {
#pragma omp parallel for num_threads(threads()) schedule(dynamic, 4)
            for (int64_t b = 0; b < nb_; ++b) {
                int bx, by, bz;
                unpack_block(pkeys[b], bx, by, bz);
                if ((((bx & 1) << 2) | ((by & 1) << 1) | (bz & 1)) != colour) continue;
                int32_t nb[8];
                neighbour_blocks(pkeys[b], nb);
                for (int64_t j = pstart[b]; j < pstart[b + 1]; ++j) p2g_particle(porder[j], nb, bx, by, bz, dx, dxinv, dt, Dinv);
            }
        }
    }

    void p2g_particle(int64_t i, const int32_t nb[8], int bx, int by, int bz, T dx, T dxinv, T dt, T Dinv) {
        const Material &mt = mats[mat[i]];
        // pars("m", pi), pars("vol", pi) (Transfer.cu:857-858): the particle's own when it carries them
        const T vol = pp_vol.empty() ? (T)mt.vol : (T)pp_vol[i];
        const T mass = pp_m.empty() ? (T)mt.rho * vol : (T)pp_m[i];
        T xp[3] = {x[i], x[n + i], x[2 * n + i]};
        T vp[3] = {v[i], v[n + i], v[2 * n + i]};
        T Cp[3][3], Fp[3][3], contrib[3][3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                Cp[r][c] = C[(3 * r + c) * n + i];
                Fp[r][c] = F[(3 * r + c) * n + i];
            }
        const T s = -dt * Dinv * vol;  // -dt * Dinv * vol (Transfer.cu:869)
        if (mt.model == MODEL_STVK_HENCKY) {
            T mu, lam;
            lame((T)mt.E, (T)mt.nu, mu, lam);
            if (!pp_mu.empty() && transfer != TRANSFER_FLIP) {  // p2g_flip has no override (Transfer.cu:892-910)
                mu = pp_mu[i];
                lam = pp_lam[i];
            }
            double Fd[3][3], tau[3][3];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) Fd[r][c] = (double)Fp[r][c];
            stvk_hencky_tau(Fd, (double)mu, (double)lam, tau);
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) contrib[r][c] = s * (T)tau[r][c];
            if (mt.aniso == ANISO_ARAP) {  // + s P_a F^T
                T Pa[3][3] = {}, mua, lama;
                lame((T)mt.E, (T)mt.nu, mua, lama);
                arap_fiber_add(Fp, fiber_of(i).data(), (T)mt.aniso_strength * mua, Pa);
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c)
                        contrib[r][c] += (s * Pa[r][0]) * Fp[c][0] + (s * Pa[r][1]) * Fp[c][1] + (s * Pa[r][2]) * Fp[c][2];
            }
        } else if (mt.model == MODEL_FCR || mt.model == MODEL_NEO_HOOKEAN || mt.model == MODEL_STABLE_NEO_HOOKEAN) {
            T mu, lam, P[3][3];
            lame((T)mt.E, (T)mt.nu, mu, lam);
            if (!pp_mu.empty() && transfer != TRANSFER_FLIP) {  // p2g_flip has no override (Transfer.cu:892-910)
                mu = pp_mu[i];
                lam = pp_lam[i];
            }
            if (mt.model == MODEL_FCR) fcr_piola(Fp, mu, lam, P);
            else neo_hookean_piola(Fp, mu, lam, mt.model == MODEL_STABLE_NEO_HOOKEAN, P);
            if (mt.aniso == ANISO_ARAP) {  // the fiber model keeps the material's E, nu (Generation.cu:63)
                T mua, lama;
                lame((T)mt.E, (T)mt.nu, mua, lama);
                arap_fiber_add(Fp, fiber_of(i).data(), (T)mt.aniso_strength * mua, P);
            }
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) {
                    T acc = T(0);
                    for (int k = 0; k < 3; ++k) acc += (s * P[r][k]) * Fp[c][k];
                    contrib[r][c] = acc;
                }
        } else {
            // weakly compressible J-fluid: Kirchhoff stress tau = K J (J - 1) I
            const T J = Fp[0][0];
            const T tau = (T)mt.bulk * J * (J - T(1));
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) contrib[r][c] = (r == c) ? s * tau : T(0);
        }
        Arena<T> a;
        make_arena(xp, dx, dxinv, a);
        if (transfer == TRANSFER_FLIP) {
            // p2g_flip (Transfer.cu:892-933): the stress term without D^-1 along the weight
            // gradient, times W once more (Transfer.cu:929); momentum without the affine term
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) contrib[r][c] /= Dinv;
            for (int ox = 0; ox < 3; ++ox)
                for (int oy = 0; oy < 3; ++oy)
                    for (int oz = 0; oz < 3; ++oz) {
                        const T W = a.w[0][ox] * a.w[1][oy] * a.w[2][oz];
                        const T g[3] = {a.dw[0][ox] * a.w[1][oy] * a.w[2][oz] * dxinv,
                                        a.w[0][ox] * a.dw[1][oy] * a.w[2][oz] * dxinv,
                                        a.w[0][ox] * a.w[1][oy] * a.dw[2][oz] * dxinv};
                        T *nd = node(nb, bx, by, bz, a.base[0] + ox, a.base[1] + oy, a.base[2] + oz);
                        T *ns = &gstar[nd - grid.data()];
                        nd[0] += W * mass;
                        for (int k = 0; k < 3; ++k) {
                            const T fdt = contrib[k][0] * g[0] + contrib[k][1] * g[1] + contrib[k][2] * g[2];
                            nd[64 * (k + 1)] += (W * mass) * vp[k];
                            ns[64 * (k + 1)] += W * fdt;
                        }
                    }
            return;
        }
        for (int ox = 0; ox < 3; ++ox)
            for (int oy = 0; oy < 3; ++oy)
                for (int oz = 0; oz < 3; ++oz) {
                    const T W = a.w[0][ox] * a.w[1][oy] * a.w[2][oz];
                    const T d[3] = {a.d[0][ox], a.d[1][oy], a.d[2][oz]};
                    T *nd = node(nb, bx, by, bz, a.base[0] + ox, a.base[1] + oy, a.base[2] + oz);
                    nd[0] += mass * W;
                    T *ns = transfer == TRANSFER_AFLIP ? &gstar[nd - grid.data()] : nullptr;
                    for (int k = 0; k < 3; ++k) {
                        const T Cd = Cp[k][0] * d[0] + Cp[k][1] * d[1] + Cp[k][2] * d[2];
                        const T fdt = contrib[k][0] * d[0] + contrib[k][1] * d[1] + contrib[k][2] * d[2];
                        if (ns) {  // p2g_aflip (Transfer.cu:977-981): momentum and stress apart
                            nd[64 * (k + 1)] += W * (mass * (vp[k] + Cd));
                            ns[64 * (k + 1)] += W * fdt;
                        } else {
                            nd[64 * (k + 1)] += W * (mass * (vp[k] + Cd) + fdt);
                        }
                    }
                }
    }

    void grid_update() {
        const T dt = (T)prm.dt;
        const T acc[3] = {(T)prm.gravity[0], (T)prm.gravity[1], (T)prm.gravity[2]};
        const int64_t G = (int64_t)gkeys.size();
        T vmax = T(0);
#pragma omp parallel for num_threads(threads()) schedule(static) reduction(max : vmax)
        for (int64_t g = 0; g < G; ++g) {
            T *blk = &grid[(size_t)g * 256];
            for (int c = 0; c < 64; ++c) {
                T mass = blk[c];
                if (mass != T(0)) {
                    mass = T(1) / mass;
                    T vel[3];
                    for (int k = 0; k < 3; ++k) vel[k] = blk[64 * (k + 1) + c] * mass;
                    if (flip_style()) {
                        // flip-style update (MPMPipeline.cu:878-895): v = mv / m kept,
                        // vstar = fs / m + v + a dt
                        T *st = &gstar[(size_t)g * 256];
                        T vs[3];
                        for (int k = 0; k < 3; ++k) {
                            blk[64 * (k + 1) + c] = vel[k];
                            vs[k] = st[64 * (k + 1) + c] * mass + vel[k] + acc[k] * dt;
                            st[64 * (k + 1) + c] = vs[k];
                        }
                        vmax = std::max(vmax, vs[0] * vs[0] + vs[1] * vs[1] + vs[2] * vs[2]);
                        continue;
                    }
                    for (int k = 0; k < 3; ++k) vel[k] += acc[k] * dt;
                    for (int k = 0; k < 3; ++k) blk[64 * (k + 1) + c] = vel[k];
                    const T vs = vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2];
                    vmax = std::max(vmax, vs);
                }
            }
        }
        maxVelSqr = vmax;
    }

    void boundary() {
        if (prm.boundary != BOUNDARY_STICKY_BOX) return;
        const int lo = prm.pad, hi = prm.n_cells - prm.pad;
        const int64_t G = (int64_t)gkeys.size();
#pragma omp parallel for num_threads(threads()) schedule(static)
        for (int64_t g = 0; g < G; ++g) {
            int bx, by, bz;
            unpack_block(gkeys[g], bx, by, bz);
            T *blk = &grid[(size_t)g * 256];
            T *vb = flip_style() ? &gstar[(size_t)g * 256] : blk;  // flip style: on vstar
            for (int c = 0; c < 64; ++c) {
                if (blk[c] == T(0)) continue;
                const int gx = 4 * bx + (c >> 4), gy = 4 * by + ((c >> 2) & 3), gz = 4 * bz + (c & 3);
                if (gx < lo || gy < lo || gz < lo || gx > hi || gy > hi || gz > hi)
                    for (int k = 1; k < 4; ++k) vb[64 * k + c] = T(0);
            }
        }
    }

    // level-set colliders on every node with m != 0, at (block key + cell) dx
    void colliders_apply() {
        if (colliders.empty()) return;
        const T dx = (T)prm.dx;
        const int64_t G = (int64_t)gkeys.size();
#pragma omp parallel for num_threads(threads()) schedule(static)
        for (int64_t g = 0; g < G; ++g) {
            int bx, by, bz;
            unpack_block(gkeys[g], bx, by, bz);
            T *blk = &grid[(size_t)g * 256];
            T *vb = flip_style() ? &gstar[(size_t)g * 256] : blk;  // flip style: on vstar
            for (int c = 0; c < 64; ++c) {
                if (blk[c] == T(0)) continue;
                const int gi[3] = {4 * bx + (c >> 4), 4 * by + ((c >> 2) & 3), 4 * bz + (c & 3)};
                double x[3];
                for (int k = 0; k < 3; ++k) x[k] = (double)((T)gi[k] * dx);
                T v[3] = {vb[64 + c], vb[128 + c], vb[192 + c]};
                for (size_t q = 0; q < colliders.size(); ++q) resolve_collision(colliders[q], leaf_index[q], x, v);
                for (int k = 0; k < 3; ++k) vb[64 * (k + 1) + c] = v[k];
            }
        }
    }

    // ZSBoundaryPrimitiveToZSGrid's p2g_momentum (Transfer.cu:1533-1578) with the linear kernel
    // (base floor(x / dx), weights 1 - f, f), the boundary grid update (MPMPipeline.cu:896-910:
    // v_b = mv / m, n_b = normalize(m nrm / m)) and ApplyGridBoundaryOnZSGrid (Boundary.cu:18-74)
    // on the nodes with mass in both grids, in FP64 sums in particle order.
    static int64_t node_key(int gx, int gy, int gz) {
        const int64_t o = 1 << 20;
        return ((int64_t)(gx + o) << 42) | ((int64_t)(gy + o) << 21) | (int64_t)(gz + o);
    }
    // the boundary particles' (m, m v, m nrm) per node, linear kernel, FP64 sums in particle order
    std::unordered_map<int64_t, std::array<double, 7>> boundary_sums() const {
        const T dxinv = T(1) / (T)prm.dx;
        auto key = node_key;
        std::unordered_map<int64_t, std::array<double, 7>> acc;
        for (int64_t p = 0; p < bnd_n; ++p) {
            int base[3];
            double w[3][2];
            for (int k = 0; k < 3; ++k) {
                const T X = (T)bnd_x[k * bnd_n + p] * dxinv;
                base[k] = (int)std::floor((double)X);
                const T f = X - (T)base[k];
                w[k][0] = (double)(T(1) - f);
                w[k][1] = (double)f;
            }
            const double m = bnd_m[p];
            for (int o = 0; o < 8; ++o) {
                const int a = o >> 2, b = (o >> 1) & 1, c = o & 1;
                const double W = w[0][a] * w[1][b] * w[2][c];
                auto &s7 = acc[key(base[0] + a, base[1] + b, base[2] + c)];
                s7[0] += W * m;
                for (int k = 0; k < 3; ++k) {
                    s7[1 + k] += W * (m * bnd_v[k * bnd_n + p]);
                    s7[4 + k] += bnd_nrm.empty() ? 0.0 : W * (m * bnd_nrm[k * bnd_n + p]);
                }
            }
        }
        return acc;
    }

    void boundary_particles_apply() {
        if (bnd_n == 0) return;
        auto key = node_key;
        const auto acc = boundary_sums();
        const int64_t G = (int64_t)gkeys.size();
        for (int64_t g = 0; g < G; ++g) {
            int bx, by, bz;
            unpack_block(gkeys[g], bx, by, bz);
            T *blk = &grid[(size_t)g * 256];
            T *vb = flip_style() ? &gstar[(size_t)g * 256] : blk;  // flip style: on vstar
            for (int c = 0; c < 64; ++c) {
                if (blk[c] == T(0)) continue;
                auto it = acc.find(key(4 * bx + (c >> 4), 4 * by + ((c >> 2) & 3), 4 * bz + (c & 3)));
                if (it == acc.end() || it->second[0] == 0.0) continue;
                const auto &s7 = it->second;
                double vo[3], nr[3];
                for (int k = 0; k < 3; ++k) {
                    vo[k] = s7[1 + k] / s7[0];
                    nr[k] = s7[4 + k] / s7[0];
                }
                if (bnd_type == COLLIDER_STICKY) {
                    for (int k = 0; k < 3; ++k) vb[64 * (k + 1) + c] = (T)vo[k];
                    continue;
                }
                const double nn = std::sqrt(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
                if (!(nn > 0.0)) continue;
                double r[3], proj = 0.0;
                for (int k = 0; k < 3; ++k) {
                    nr[k] /= nn;
                    r[k] = (double)vb[64 * (k + 1) + c] - vo[k];
                    proj += nr[k] * r[k];
                }
                if (bnd_type == COLLIDER_SLIP || proj < 0.0)
                    for (int k = 0; k < 3; ++k) r[k] -= proj * nr[k];
                for (int k = 0; k < 3; ++k) vb[64 * (k + 1) + c] = (T)(r[k] + vo[k]);
            }
        }
    }

    // StepZSBoundary (Generation.cu:1663-1672)
    void colliders_step(double dt) {
        for (Collider &c : colliders) {
            for (int k = 0; k < 3; ++k) c.translation[k] += c.translation_rate[k] * (float)dt;
            c.scale += c.scale_rate * (float)dt;
        }
    }

    void g2p() {
        const T dx = (T)prm.dx, dxinv = T(1) / dx, dt = (T)prm.dt;
        const T Dinv = T(4) * dxinv * dxinv;
        const int64_t nb_ = (int64_t)pkeys.size();
#pragma omp parallel for num_threads(threads()) schedule(dynamic, 4)
        for (int64_t b = 0; b < nb_; ++b) {
            int bx, by, bz;
            unpack_block(pkeys[b], bx, by, bz);
            int32_t nb[8];
            neighbour_blocks(pkeys[b], nb);
            for (int64_t j = pstart[b]; j < pstart[b + 1]; ++j) {
                const int64_t i = porder[j];
                T xp[3] = {x[i], x[n + i], x[2 * n + i]};
                Arena<T> a;
                make_arena(xp, dx, dxinv, a);
                const bool aflip = flip_style(), nonaffine = transfer == TRANSFER_FLIP;
                T vel[3] = {0, 0, 0}, vst[3] = {0, 0, 0}, Cn[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                for (int ox = 0; ox < 3; ++ox)
                    for (int oy = 0; oy < 3; ++oy)
                        for (int oz = 0; oz < 3; ++oz) {
                            const T W = a.w[0][ox] * a.w[1][oy] * a.w[2][oz];
                            const T d[3] = {a.d[0][ox], a.d[1][oy], a.d[2][oz]};
                            const T *nd = node(nb, bx, by, bz, a.base[0] + ox, a.base[1] + oy, a.base[2] + oz);
                            const T vi[3] = {nd[64], nd[128], nd[192]};
                            for (int k = 0; k < 3; ++k) vel[k] += vi[k] * W;
                            // AFLIP: C from vstar (Transfer.cu:1196)
                            const T *ns = aflip ? &gstar[nd - grid.data()] : nd;
                            const T si[3] = {ns[64], ns[128], ns[192]};
                            for (int k = 0; k < 3; ++k) vst[k] += si[k] * W;
                            if (nonaffine) {
                                // velocity gradient sum vstar_i grad W^T (Transfer.cu:1232-1240)
                                const T g[3] = {a.dw[0][ox] * a.w[1][oy] * a.w[2][oz] * dxinv,
                                                a.w[0][ox] * a.dw[1][oy] * a.w[2][oz] * dxinv,
                                                a.w[0][ox] * a.w[1][oy] * a.dw[2][oz] * dxinv};
                                for (int r = 0; r < 3; ++r)
                                    for (int c = 0; c < 3; ++c) Cn[r][c] += si[r] * g[c];
                                continue;
                            }
                            const T WD = W * Dinv;
                            for (int r = 0; r < 3; ++r)
                                for (int c = 0; c < 3; ++c) Cn[r][c] += WD * (si[r] * d[c]);
                        }
                if (aflip) {
                    // v = vstar + flip (v0 - v), x += dt (vstar + beta flip (v0 - v)) (Transfer.cu:1199-1206)
                    // FLIP (Transfer.cu:1244-1247): x += vstar dt, the beta term absent
                    const T fl = (T)kFlipBlend, bt = nonaffine ? T(0) : beta[i];
                    for (int k = 0; k < 3; ++k) {
                        const T dv = v[k * n + i] - vel[k];
                        v[k * n + i] = vst[k] + fl * dv;
                        x[k * n + i] = xp[k] + dt * (vst[k] + bt * fl * dv);
                    }
                } else {
                    for (int k = 0; k < 3; ++k) {
                        v[k * n + i] = vel[k];
                        x[k * n + i] = xp[k] + vel[k] * dt;
                    }
                }
                if (!nonaffine)  // FLIP keeps the particle's C
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c) C[(3 * r + c) * n + i] = Cn[r][c];
                const Material &mt = mats[mat[i]];
                if (mt.model != MODEL_JFLUID) {
                    T Fo[3][3], Fn[3][3];
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c) Fo[r][c] = F[(3 * r + c) * n + i];
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c) {
                            T acc = T(0);
                            for (int k = 0; k < 3; ++k) acc += ((r == k ? T(1) : T(0)) + Cn[r][k] * dt) * Fo[k][c];
                            Fn[r][c] = acc;
                        }
                    if (mt.plasticity != PLASTIC_NONE) {  // ZSReturnMapping after G2P
                        T mu, lam;
                        lame((T)mt.E, (T)mt.nu, mu, lam);
                        double Fd[3][3];
                        for (int r = 0; r < 3; ++r)
                            for (int c = 0; c < 3; ++c) Fd[r][c] = (double)Fn[r][c];
                        int changed;
                        if (mt.plasticity == PLASTIC_CAM_CLAY) {
                            double lj = (double)logJp[i];
                            changed = project_cam_clay(Fd, (double)mu, (double)lam, (double)mt.friction_angle,
                                                       (double)mt.nacc_beta, (double)mt.nacc_xi, lj);
                            if (changed) logJp[i] = (T)lj;
                        } else {
                            changed = project_strain(Fd, (double)mu, (double)lam, mt.plasticity,
                                                     dp_alpha(mt.friction_angle), (double)mt.yield_stress);
                        }
                        if (changed)
                            for (int r = 0; r < 3; ++r)
                                for (int c = 0; c < 3; ++c) Fn[r][c] = (T)Fd[r][c];
                    }
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c) F[(3 * r + c) * n + i] = Fn[r][c];
                } else {
                    const T tr = Cn[0][0] + Cn[1][1] + Cn[2][2];
                    F[i] = (T(1) + dt * tr) * F[i];
                }
            }
        }
    }

    // ComputeParticleBeta with a level-set boundary (Boundary.cu:284-313): at x + dt v,
    // inside a collider (phi < 0) and (v - dbdt) . n < 0 -> beta 0 (the level set has no
    // material velocity); else beta 0.5 if det F (FP64; J for J-fluid) > jpc, else 0.
    // With boundary particles, also the boundary-grid branch (Boundary.cu:173-219): the 8 nodes of
    // the linear stencil at x + dt v (those in the partition's grid blocks) with boundary mass and
    // (v - v_b) . n_b < 0 make the particle moving in.
    void compute_particle_beta(double h, double jpc) {
        std::unordered_map<int64_t, std::array<double, 7>> acc;
        if (bnd_n > 0) {
            partition();
            acc = boundary_sums();
        }
        const T dxinv = T(1) / (T)prm.dx;
#pragma omp parallel for num_threads(threads()) schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            double pos[3], vp[3];
            for (int k = 0; k < 3; ++k) {
                vp[k] = (double)v[k * n + i];
                pos[k] = (double)(x[k * n + i] + (T)h * v[k * n + i]);
            }
            bool moving_in = false;
            for (size_t q = 0; q < colliders.size(); ++q) {
                const Collider &c = colliders[q];
                double R[3][3], X[3], g[3];
                collider_rotation(c, R);
                for (int a = 0; a < 3; ++a) {
                    X[a] = 0.0;
                    for (int b = 0; b < 3; ++b) X[a] += R[b][a] * (pos[b] - (double)c.translation[b]);
                    X[a] /= (double)c.scale;
                }
                if (!(shape_phi(c, leaf_index[q], X, g) < 0.0)) continue;
                double dot = 0.0;
                for (int a = 0; a < 3; ++a) {
                    double na = 0.0;
                    for (int b = 0; b < 3; ++b) na += R[a][b] * g[b];
                    dot += (vp[a] - (double)c.translation_rate[a]) * na;
                }
                if (dot < 0.0) {
                    moving_in = true;
                    break;
                }
            }
            if (!moving_in && bnd_n > 0) {
                int base[3];
                for (int k = 0; k < 3; ++k) base[k] = (int)std::floor((double)((T)pos[k] * dxinv));
                for (int o = 0; o < 8 && !moving_in; ++o) {
                    const int g[3] = {base[0] + (o >> 2), base[1] + ((o >> 1) & 1), base[2] + (o & 1)};
                    if (gmap.find(pack_block(fdiv4(g[0]), fdiv4(g[1]), fdiv4(g[2]))) == gmap.end()) continue;
                    auto it = acc.find(node_key(g[0], g[1], g[2]));
                    if (it == acc.end() || it->second[0] == 0.0) continue;
                    double dot = 0.0, nn = 0.0;
                    for (int k = 0; k < 3; ++k) {
                        const double nk = it->second[4 + k] / it->second[0];
                        dot += (vp[k] - it->second[1 + k] / it->second[0]) * nk;
                        nn += nk * nk;
                    }
                    moving_in = nn > 0.0 && dot < 0.0;
                }
            }
            if (moving_in) {
                beta[i] = T(0);
                continue;
            }
            double J;
            if (mats[mat[i]].model == MODEL_JFLUID) {
                J = (double)F[i];
            } else {
                double Fd[3][3];
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) Fd[r][c] = (double)F[(3 * r + c) * n + i];
                J = det3(Fd);
            }
            // jpc < 0: JpcDefault = 1 for every material.  The reference sets 1.1 only when its
            // anisotropic-model variant holds a Cam-clay model (Boundary.cu:389-391, 412-414,
            // 466-468), which that variant (CUDA/Structures.hpp:30) cannot hold.
            const double jp = jpc >= 0.0 ? jpc : 1.0;
            beta[i] = J > jp ? T(0.5) : T(0);
        }
    }

    void substep() {
        partition();
        p2g();
        grid_update();
        boundary();
        colliders_apply();
        boundary_particles_apply();
        colliders_step(prm.dt);
        g2p();
    }
};

}  // namespace orc

// ---------------------------------------------------------------- C ABI
using namespace orc;

#define ORC_API extern "C" __attribute__((visibility("default")))

template <typename T>
static void *create_t(const Params *prm, const Material *mats, int32_t nmat, int64_t n, const T *x, const T *v,
                      const T *C, const T *F, const int32_t *mat) {
    auto *s = new Session<T>();
    s->prm = *prm;
    s->mats.assign(mats, mats + nmat);
    s->n = n;
    s->x.assign(x, x + 3 * n);
    s->v.assign(v, v + 3 * n);
    s->C.assign(C, C + 9 * n);
    s->F.assign(F, F + 9 * n);
    s->mat.assign(mat, mat + n);
    s->beta.resize(n);
    s->logJp.assign(n, (T)kLogJpInit);
    for (int64_t i = 0; i < n; ++i) s->beta[i] = (T)s->mats[mat[i]].flip_beta;
    return s;
}

template <typename T> static void aux_t(void *h, T *beta, T *logJp, int32_t set) {
    auto *s = (Session<T> *)h;
    if (set) {
        if (beta) std::copy(beta, beta + s->n, s->beta.begin());
        if (logJp) std::copy(logJp, logJp + s->n, s->logJp.begin());
    } else {
        if (beta) std::copy(s->beta.begin(), s->beta.end(), beta);
        if (logJp) std::copy(s->logJp.begin(), s->logJp.end(), logJp);
    }
}

template <typename T> static void get_t(void *h, T *x, T *v, T *C, T *F) {
    auto *s = (Session<T> *)h;
    std::copy(s->x.begin(), s->x.end(), x);
    std::copy(s->v.begin(), s->v.end(), v);
    std::copy(s->C.begin(), s->C.end(), C);
    std::copy(s->F.begin(), s->F.end(), F);
}

template <typename T> static int32_t stage_t(void *h, int32_t stage) {
    auto *s = (Session<T> *)h;
    switch (stage) {
    case 0: s->partition(); break;
    case 1: s->p2g(); break;
    case 2: s->grid_update(); break;
    case 3:
        s->boundary();
        s->colliders_apply();
        s->boundary_particles_apply();
        break;
    case 4: s->g2p(); break;
    default: return 1;
    }
    return 0;
}

template <typename T> static void grid_t(void *h, int32_t *keys, T *data) {
    auto *s = (Session<T> *)h;
    for (size_t g = 0; g < s->gkeys.size(); ++g) {
        int bx, by, bz;
        unpack_block(s->gkeys[g], bx, by, bz);
        keys[3 * g] = 4 * bx;
        keys[3 * g + 1] = 4 * by;
        keys[3 * g + 2] = 4 * bz;
    }
    std::copy(s->grid.begin(), s->grid.end(), data);
}

ORC_API void *orc_create_f32(const Params *p, const Material *m, int32_t nm, int64_t n, const float *x,
                             const float *v, const float *C, const float *F, const int32_t *mat) {
    return create_t<float>(p, m, nm, n, x, v, C, F, mat);
}
ORC_API void *orc_create_f64(const Params *p, const Material *m, int32_t nm, int64_t n, const double *x,
                             const double *v, const double *C, const double *F, const int32_t *mat) {
    return create_t<double>(p, m, nm, n, x, v, C, F, mat);
}
ORC_API void orc_destroy_f32(void *h) { delete (Session<float> *)h; }
ORC_API void orc_destroy_f64(void *h) { delete (Session<double> *)h; }
ORC_API void orc_set_dt_f32(void *h, double dt) { ((Session<float> *)h)->prm.dt = dt; }
ORC_API void orc_set_dt_f64(void *h, double dt) { ((Session<double> *)h)->prm.dt = dt; }
ORC_API void orc_set_threads_f32(void *h, int32_t t) { ((Session<float> *)h)->prm.threads = t; }
ORC_API void orc_set_threads_f64(void *h, int32_t t) { ((Session<double> *)h)->prm.threads = t; }
ORC_API int32_t orc_stage_f32(void *h, int32_t st) { return stage_t<float>(h, st); }
ORC_API int32_t orc_stage_f64(void *h, int32_t st) { return stage_t<double>(h, st); }
ORC_API void orc_substeps_f32(void *h, int32_t k) {
    for (int i = 0; i < k; ++i) ((Session<float> *)h)->substep();
}
ORC_API void orc_substeps_f64(void *h, int32_t k) {
    for (int i = 0; i < k; ++i) ((Session<double> *)h)->substep();
}
// boundary particles: x, v, nrm (3, n) and m (n) float32 SoA (nrm may be NULL), response type
template <typename T>
static void set_bnd_t(void *h, int64_t n, const float *x, const float *v, const float *m, const float *nrm,
                      int32_t type) {
    auto *s = (Session<T> *)h;
    s->bnd_n = n;
    s->bnd_type = type;
    s->bnd_x.assign(x, x + 3 * n);
    s->bnd_v.assign(v, v + 3 * n);
    s->bnd_m.assign(m, m + n);
    if (nrm) s->bnd_nrm.assign(nrm, nrm + 3 * n);
    else s->bnd_nrm.clear();
}
ORC_API void orc_set_boundary_particles_f32(void *h, int64_t n, const float *x, const float *v, const float *m,
                                            const float *nrm, int32_t type) {
    set_bnd_t<float>(h, n, x, v, m, nrm, type);
}
ORC_API void orc_set_boundary_particles_f64(void *h, int64_t n, const float *x, const float *v, const float *m,
                                            const float *nrm, int32_t type) {
    set_bnd_t<double>(h, n, x, v, m, nrm, type);
}
// per-particle beta / logJp: set (set = 1) or get (set = 0); either pointer may be NULL
ORC_API void orc_aux_f32(void *h, float *beta, float *logJp, int32_t set) { aux_t<float>(h, beta, logJp, set); }
ORC_API void orc_aux_f64(void *h, double *beta, double *logJp, int32_t set) { aux_t<double>(h, beta, logJp, set); }
// per-particle Lame parameters (float32, n each); NULL removes them
template <typename T> static void lame_t(void *h, const float *mu, const float *lam) {
    auto *s = (Session<T> *)h;
    if (mu && lam) {
        s->pp_mu.assign(mu, mu + s->n);
        s->pp_lam.assign(lam, lam + s->n);
    } else {
        s->pp_mu.clear();
        s->pp_lam.clear();
    }
}
// per-particle mass and volume (float32, n each); NULL removes them
template <typename T> static void mass_vol_t(void *h, const float *m, const float *vol) {
    auto *s = (Session<T> *)h;
    if (m && vol) {
        s->pp_m.assign(m, m + s->n);
        s->pp_vol.assign(vol, vol + s->n);
    } else {
        s->pp_m.clear();
        s->pp_vol.clear();
    }
}
ORC_API void orc_set_mass_vol_f32(void *h, const float *m, const float *vol) { mass_vol_t<float>(h, m, vol); }
ORC_API void orc_set_mass_vol_f64(void *h, const float *m, const float *vol) { mass_vol_t<double>(h, m, vol); }
// per-particle fiber directions (float32 SoA 3 x n); NULL: (0, 1, 0) for all
template <typename T> static void fiber_t(void *h, const float *a) {
    auto *s = (Session<T> *)h;
    if (a) s->fiber.assign(a, a + 3 * s->n);
    else s->fiber.clear();
}
ORC_API void orc_set_fiber_f32(void *h, const float *a) { fiber_t<float>(h, a); }
ORC_API void orc_set_fiber_f64(void *h, const float *a) { fiber_t<double>(h, a); }
ORC_API void orc_arap_fiber_piola_f64(const double *F9, const double *a, double mu_a, double *P9) {
    double F[3][3], P[3][3] = {};
    for (int i = 0; i < 9; ++i) F[i / 3][i % 3] = F9[i];
    arap_fiber_add<double>(F, a, mu_a, P);
    for (int i = 0; i < 9; ++i) P9[i] = P[i / 3][i % 3];
}
ORC_API void orc_set_lame_f32(void *h, const float *mu, const float *lam) { lame_t<float>(h, mu, lam); }
ORC_API void orc_set_lame_f64(void *h, const float *mu, const float *lam) { lame_t<double>(h, mu, lam); }
ORC_API void orc_compute_particle_beta_f32(void *h, double dt, double jpc) {
    ((Session<float> *)h)->compute_particle_beta(dt, jpc);
}
ORC_API void orc_compute_particle_beta_f64(void *h, double dt, double jpc) {
    ((Session<double> *)h)->compute_particle_beta(dt, jpc);
}
ORC_API void orc_set_transfer_f32(void *h, int32_t t) { ((Session<float> *)h)->transfer = t; }
ORC_API void orc_set_transfer_f64(void *h, int32_t t) { ((Session<double> *)h)->transfer = t; }
ORC_API void orc_set_colliders_f32(void *h, const Collider *c, int32_t n) {
    ((Session<float> *)h)->set_colliders(c, n);
}
ORC_API void orc_set_colliders_f64(void *h, const Collider *c, int32_t n) {
    ((Session<double> *)h)->set_colliders(c, n);
}
// one collider response at x on v (FP64); returns 1 if x is inside
ORC_API int32_t orc_resolve_collision_f64(const Collider *c, const double *x, double *v) {
    return resolve_collision<double>(*c, LeafIndex(*c), x, v);
}
// the collider's shape alone: phi and its gradient at X in the collider frame (FP64)
ORC_API double orc_shape_phi_f64(const Collider *c, const double *X, double *g) {
    return shape_phi(*c, LeafIndex(*c), X, g);
}
ORC_API double orc_max_vel_sqr_f32(void *h) { return ((Session<float> *)h)->maxVelSqr; }
ORC_API double orc_max_vel_sqr_f64(void *h) { return ((Session<double> *)h)->maxVelSqr; }
ORC_API void orc_get_f32(void *h, float *x, float *v, float *C, float *F) { get_t<float>(h, x, v, C, F); }
ORC_API void orc_get_f64(void *h, double *x, double *v, double *C, double *F) { get_t<double>(h, x, v, C, F); }
ORC_API int64_t orc_num_pblocks_f32(void *h) { return (int64_t)((Session<float> *)h)->pkeys.size(); }
ORC_API int64_t orc_num_gblocks_f32(void *h) { return (int64_t)((Session<float> *)h)->gkeys.size(); }
ORC_API int64_t orc_num_gblocks_f64(void *h) { return (int64_t)((Session<double> *)h)->gkeys.size(); }
ORC_API void orc_grid_f32(void *h, int32_t *keys, float *data) { grid_t<float>(h, keys, data); }
// the FLIP transfers' vstar grid (same block order and layout as orc_grid; empty under APIC)
ORC_API void orc_gstar_f64(void *h, double *data) {
    auto *s = (Session<double> *)h;
    std::copy(s->gstar.begin(), s->gstar.end(), data);
}
ORC_API void orc_grid_f64(void *h, int32_t *keys, double *data) { grid_t<double>(h, keys, data); }
// particle order grouped by block (stable by particle index), for binning parity
ORC_API void orc_porder_f32(void *h, int64_t *order) {
    auto *s = (Session<float> *)h;
    std::copy(s->porder.begin(), s->porder.end(), order);
}

// ---- primitives, exposed for closed-form tests
ORC_API void orc_svd3(const double *F9, double *U9, double *s3, double *V9) {
    double F[3][3], U[3][3], V[3][3];
    for (int i = 0; i < 9; ++i) F[i / 3][i % 3] = F9[i];
    svd3(F, U, s3, V);
    for (int i = 0; i < 9; ++i) {
        U9[i] = U[i / 3][i % 3];
        V9[i] = V[i / 3][i % 3];
    }
}
ORC_API void orc_polar_rotation(const double *F9, double *R9) {
    double F[3][3], R[3][3];
    for (int i = 0; i < 9; ++i) F[i / 3][i % 3] = F9[i];
    polar_rotation(F, R);
    for (int i = 0; i < 9; ++i) R9[i] = R[i / 3][i % 3];
}
ORC_API void orc_fcr_piola_f64(const double *F9, double mu, double lam, double *P9) {
    double F[3][3], P[3][3];
    for (int i = 0; i < 9; ++i) F[i / 3][i % 3] = F9[i];
    fcr_piola<double>(F, mu, lam, P);
    for (int i = 0; i < 9; ++i) P9[i] = P[i / 3][i % 3];
}
ORC_API void orc_fcr_piola_f32(const float *F9, float mu, float lam, float *P9) {
    float F[3][3], P[3][3];
    for (int i = 0; i < 9; ++i) F[i / 3][i % 3] = F9[i];
    fcr_piola<float>(F, mu, lam, P);
    for (int i = 0; i < 9; ++i) P9[i] = P[i / 3][i % 3];
}
ORC_API void orc_lame_f32(float E, float nu, float *mu, float *lam) { lame<float>(E, nu, *mu, *lam); }
ORC_API void orc_neo_hookean_piola_f64(const double *F9, double mu, double lam, int32_t stable, double *P9) {
    double F[3][3], P[3][3];
    for (int i = 0; i < 9; ++i) F[i / 3][i % 3] = F9[i];
    neo_hookean_piola<double>(F, mu, lam, stable != 0, P);
    for (int i = 0; i < 9; ++i) P9[i] = P[i / 3][i % 3];
}
ORC_API void orc_stvk_hencky_tau_f64(const double *F9, double mu, double lam, double *tau9) {
    double F[3][3], tau[3][3];
    for (int i = 0; i < 9; ++i) F[i / 3][i % 3] = F9[i];
    stvk_hencky_tau(F, mu, lam, tau);
    for (int i = 0; i < 9; ++i) tau9[i] = tau[i / 3][i % 3];
}
// return mapping of one F (in place); returns 1 if it changed
ORC_API int32_t orc_project_strain_f64(double *F9, double mu, double lam, int32_t plasticity, double friction_angle,
                                       double yield_stress) {
    double F[3][3];
    for (int i = 0; i < 9; ++i) F[i / 3][i % 3] = F9[i];
    const int r = project_strain(F, mu, lam, plasticity, dp_alpha(friction_angle), yield_stress);
    for (int i = 0; i < 9; ++i) F9[i] = F[i / 3][i % 3];
    return r;
}
// Cam-clay return mapping of one F (in place) and its logJp; returns 1 if F changed
ORC_API int32_t orc_project_cam_clay_f64(double *F9, double mu, double lam, double friction_angle, double beta,
                                         double xi, double *logJp) {
    double F[3][3];
    for (int i = 0; i < 9; ++i) F[i / 3][i % 3] = F9[i];
    const int r = project_cam_clay(F, mu, lam, friction_angle, beta, xi, *logJp);
    for (int i = 0; i < 9; ++i) F9[i] = F[i / 3][i % 3];
    return r;
}
// arena weights/diffs of one particle: base[3], w[9] (axis-major), d[9]
ORC_API void orc_arena_f64(const double *xp, double dx, int32_t *base, double *w, double *d) {
    Arena<double> a;
    make_arena<double>(xp, dx, 1.0 / dx, a);
    for (int k = 0; k < 3; ++k) {
        base[k] = a.base[k];
        for (int o = 0; o < 3; ++o) {
            w[3 * k + o] = a.w[k][o];
            d[3 * k + o] = a.d[k][o];
        }
    }
}
// the 27 stencil weights W and weight gradients grad W (d/dx_p, FP64) of one
// particle, node o = (ox, oy, oz) at 9 ox + 3 oy + oz: W[27], grad[27 * 3]
ORC_API void orc_weight_grad_f64(const double *xp, double dx, int32_t *base, double *W, double *grad) {
    Arena<double> a;
    make_arena<double>(xp, dx, 1.0 / dx, a);
    for (int k = 0; k < 3; ++k) base[k] = a.base[k];
    for (int ox = 0; ox < 3; ++ox)
        for (int oy = 0; oy < 3; ++oy)
            for (int oz = 0; oz < 3; ++oz) {
                const int j = 9 * ox + 3 * oy + oz;
                W[j] = a.w[0][ox] * a.w[1][oy] * a.w[2][oz];
                grad[3 * j] = a.dw[0][ox] * a.w[1][oy] * a.w[2][oz] / dx;
                grad[3 * j + 1] = a.w[0][ox] * a.dw[1][oy] * a.w[2][oz] / dx;
                grad[3 * j + 2] = a.w[0][ox] * a.w[1][oy] * a.dw[2][oz] / dx;
            }
}
// cell coordinates floor(x*dxinv - 0.5) with FP32 x*dxinv (MPMPipeline.cu:77-83)
ORC_API void orc_cell_coords_f32(const float *x, int64_t n, float dx, int32_t *coords) {
    const float dxinv = 1.f / dx;
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) coords[k * n + i] = block_floor(x[k * n + i] * dxinv);
}
