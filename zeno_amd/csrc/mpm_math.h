// Device math for the MPM kernels: quadratic B-spline arena, FP64 signed SVD /
// polar rotation, fixed-corotated stress.  Compiled with -ffp-contract=off:
// every product and sum rounds as written, fused multiply-adds appear only
// where fmaf() is written out, so two kernels evaluating the same helper on
// the same inputs produce bit-identical results.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/zmpm.h"

namespace zmpm {

// Quadratic B-spline weights of one particle (zpc make_local_arena, quadratic,
// collocated, as called at projects/CuLagrange/mpm/Transfer.cu:870):
// X = x / dx, base = floor(X - 0.5), f = X - base in [0.5, 1.5),
// w0 = (1.5-f)^2/2, w1 = 0.75-(f-1)^2, w2 = (f-0.5)^2/2.
struct Arena {
    int base[3];
    float f[3];
    float w[3][3];
};

// floor(x * (1/dx) - 0.5) with an FP32 product and an exact FP64 subtraction,
// the key arithmetic of projects/CuLagrange/mpm/MPMPipeline.cu:77-83.
__device__ __forceinline__ int cell_floor(float X) { return (int)floor((double)X - 0.5); }

__device__ __forceinline__ void make_arena(const float xp[3], float dxinv, Arena &a) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float X = xp[k] * dxinv;
        a.base[k] = cell_floor(X);
        const float f = X - (float)a.base[k];
        a.f[k] = f;
        a.w[k][0] = 0.5f * (1.5f - f) * (1.5f - f);
        a.w[k][1] = 0.75f - (f - 1.0f) * (f - 1.0f);
        a.w[k][2] = 0.5f * (f - 0.5f) * (f - 0.5f);
    }
}

// Symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations (Golub & Van
// Loan, "Matrix Computations", 4th ed., sec. 8.5.2-8.5.3).
__device__ __forceinline__ void jacobi_eig3(double A[3][3], double V[3][3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) V[i][j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 16; ++sweep) {
        const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        const double dia = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
        if (off <= 1e-34 * dia || off == 0.0) break;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int p = r == 2 ? 1 : 0, q = r == 0 ? 1 : 2, k = 3 - p - q;
            const double apq = A[p][q];
            if (apq == 0.0) continue;
            const double tau = (A[q][q] - A[p][p]) / (2.0 * apq);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
            A[p][p] -= t * apq;
            A[q][q] += t * apq;
            A[p][q] = A[q][p] = 0.0;
            const double akp = A[k][p], akq = A[k][q];
            A[k][p] = A[p][k] = c * akp - s * akq;
            A[k][q] = A[q][k] = s * akp + c * akq;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const double vp = V[i][p], vq = V[i][q];
                V[i][p] = c * vp - s * vq;
                V[i][q] = s * vp + c * vq;
            }
        }
    }
}

// Rotation factor of the signed SVD F = U diag(s) V^T (U, V proper rotations,
// s sorted by magnitude, the smallest carrying sign(det F)): V from the
// eigenvectors of F^T F, U from a Givens QR of F V (McAdams et al. 2011,
// UW-Madison TR1690, with exact FP64 rotations; inversion convention of
// Irving, Teran, Fedkiw 2004).  Returns R = U V^T.
__device__ __forceinline__ void polar_rotation(const double F[3][3], double R[3][3]) {
    double S[3][3], V[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) S[i][j] = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
    jacobi_eig3(S, V);
    double ev[3] = {S[0][0], S[1][1], S[2][2]};
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2 - i; ++j)
            if (ev[j] < ev[j + 1]) {
                const double te = ev[j];
                ev[j] = ev[j + 1];
                ev[j + 1] = te;
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    const double tv = V[r][j];
                    V[r][j] = V[r][j + 1];
                    V[r][j + 1] = tv;
                }
            }
    const double detV = V[0][0] * (V[1][1] * V[2][2] - V[1][2] * V[2][1]) -
                        V[0][1] * (V[1][0] * V[2][2] - V[1][2] * V[2][0]) +
                        V[0][2] * (V[1][0] * V[2][1] - V[1][1] * V[2][0]);
    if (detV < 0.0) {
#pragma unroll
        for (int r = 0; r < 3; ++r) V[r][2] = -V[r][2];
    }
    double B[3][3], U[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            B[i][j] = F[i][0] * V[0][j] + F[i][1] * V[1][j] + F[i][2] * V[2][j];
            U[i][j] = (i == j) ? 1.0 : 0.0;
        }
#pragma unroll
    for (int g = 0; g < 3; ++g) {
        const int p = g == 2 ? 1 : 0, q = g == 0 ? 1 : 2, col = g == 2 ? 1 : 0;
        const double a = B[p][col], b = B[q][col];
        const double r = sqrt(a * a + b * b);
        double c = 1.0, sn = 0.0;
        if (r > 0.0) {
            c = a / r;
            sn = b / r;
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double bp = B[p][j], bq = B[q][j];
            B[p][j] = c * bp + sn * bq;
            B[q][j] = -sn * bp + c * bq;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double up = U[i][p], uq = U[i][q];
            U[i][p] = c * up + sn * uq;
            U[i][q] = -sn * up + c * uq;
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) R[i][j] = U[i][0] * V[j][0] + U[i][1] * V[j][1] + U[i][2] * V[j][2];
}

// Rotation factor of F for det F > 0 by the scaled Newton iteration
// X <- (g X + X^-T / g) / 2 (Higham, "Computing the polar decomposition - with
// applications", SIAM J. Sci. Stat. Comput. 7, 1986), Frobenius-norm scaling
// g = sqrt(|X^-1| / |X|) until the step is below 1e-2, unscaled after, in FP64.
// For det F > 0 the polar factor is the rotation of the signed SVD, so this
// agrees with polar_rotation to FP64 rounding.  Returns false (R untouched) if
// det F is not clearly positive or the iteration has not converged in 16
// steps; the caller then uses polar_rotation.
__device__ __forceinline__ bool polar_rotation_newton(const double F[3][3], double R[3][3]) {
    double X[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) X[i][j] = F[i][j];
    bool scaled = true;
    for (int it = 0; it < 16; ++it) {
        double cof[3][3];
        cof[0][0] = X[1][1] * X[2][2] - X[1][2] * X[2][1];
        cof[0][1] = X[1][2] * X[2][0] - X[1][0] * X[2][2];
        cof[0][2] = X[1][0] * X[2][1] - X[1][1] * X[2][0];
        cof[1][0] = X[0][2] * X[2][1] - X[0][1] * X[2][2];
        cof[1][1] = X[0][0] * X[2][2] - X[0][2] * X[2][0];
        cof[1][2] = X[0][1] * X[2][0] - X[0][0] * X[2][1];
        cof[2][0] = X[0][1] * X[1][2] - X[0][2] * X[1][1];
        cof[2][1] = X[0][2] * X[1][0] - X[0][0] * X[1][2];
        cof[2][2] = X[0][0] * X[1][1] - X[0][1] * X[1][0];
        const double det = X[0][0] * cof[0][0] + X[0][1] * cof[0][1] + X[0][2] * cof[0][2];
        double nx = 0.0, nc = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                nx += X[i][j] * X[i][j];
                nc += cof[i][j] * cof[i][j];
            }
        if (!(det > 1e-12 * nx * sqrt(nx))) return false;
        // X^-T = cof / det; |X^-1|_F = |cof|_F / det
        const double g = scaled ? sqrt(sqrt(nc / nx) / det) : 1.0;
        const double a = 0.5 * g, b = 0.5 / (g * det);
        double step = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double xn = a * X[i][j] + b * cof[i][j];
                step += (xn - X[i][j]) * (xn - X[i][j]);
                X[i][j] = xn;
            }
        if (step < 1e-4) scaled = false;
        if (step < 1e-26) {
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) R[i][j] = X[i][j];
            return true;
        }
    }
    return false;
}

// Compressible Neo-Hookean ("nhk"; zpc's NeoHookean is absent: the energy of
// Jiang et al. 2016, "The Material Point Method for Simulating Continuum
// Materials", eq. 48): psi = mu/2 (tr F^T F - 3) - mu log J + lam/2 (log J)^2,
// P = mu F + (lam log J - mu) cof(F) / J; J is clamped to >= 1e-6 so an
// inverted F keeps a finite (strongly restoring) stress.
__device__ __forceinline__ void neo_hookean_piola(const float F[3][3], float mu, float lam, float P[3][3]) {
    float cof[3][3];
    cof[0][0] = F[1][1] * F[2][2] - F[1][2] * F[2][1];
    cof[0][1] = F[1][2] * F[2][0] - F[1][0] * F[2][2];
    cof[0][2] = F[1][0] * F[2][1] - F[1][1] * F[2][0];
    cof[1][0] = F[0][2] * F[2][1] - F[0][1] * F[2][2];
    cof[1][1] = F[0][0] * F[2][2] - F[0][2] * F[2][0];
    cof[1][2] = F[0][1] * F[2][0] - F[0][0] * F[2][1];
    cof[2][0] = F[0][1] * F[1][2] - F[0][2] * F[1][1];
    cof[2][1] = F[0][2] * F[1][0] - F[0][0] * F[1][2];
    cof[2][2] = F[0][0] * F[1][1] - F[0][1] * F[1][0];
    const float J = fmaxf(F[0][0] * cof[0][0] + F[0][1] * cof[0][1] + F[0][2] * cof[0][2], 1e-6f);
    const float k = (lam * logf(J) - mu) / J;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) P[i][j] = mu * F[i][j] + k * cof[i][j];
}

// Stable Neo-Hookean in its invariant form ("snhk"; zpc's
// StableNeohookeanInvarient is absent: Smith, de Goes, Kim 2018, "Stable
// Neo-Hookean Flesh Simulation", the energy without the log term):
// psi = mu/2 (I_C - 3) - mu (J - 1) + lam/2 (J - 1)^2, P = mu F + (lam (J - 1) - mu) cof(F).
// Rest-stable (P(I) = 0) and defined for inverted F.
__device__ __forceinline__ void stable_neo_hookean_piola(const float F[3][3], float mu, float lam, float P[3][3]) {
    float cof[3][3];
    cof[0][0] = F[1][1] * F[2][2] - F[1][2] * F[2][1];
    cof[0][1] = F[1][2] * F[2][0] - F[1][0] * F[2][2];
    cof[0][2] = F[1][0] * F[2][1] - F[1][1] * F[2][0];
    cof[1][0] = F[0][2] * F[2][1] - F[0][1] * F[2][2];
    cof[1][1] = F[0][0] * F[2][2] - F[0][2] * F[2][0];
    cof[1][2] = F[0][1] * F[2][0] - F[0][0] * F[2][1];
    cof[2][0] = F[0][1] * F[1][2] - F[0][2] * F[1][1];
    cof[2][1] = F[0][2] * F[1][0] - F[0][0] * F[1][2];
    cof[2][2] = F[0][0] * F[1][1] - F[0][1] * F[1][0];
    const float J = F[0][0] * cof[0][0] + F[0][1] * cof[0][1] + F[0][2] * cof[0][2];
    const float k = lam * (J - 1.0f) - mu;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) P[i][j] = mu * F[i][j] + k * cof[i][j];
}

// Anisotropic ARAP fiber term added to P ("arap", AnisotropicArap{E, nu,
// strength}, Generation.cu:60-63, Transfer.cu:866-867; zpc's model is absent):
// the energy of Kim, De Goes, Iben 2019, "Anisotropic Elasticity for
// Inversion-Safety and Element Rehabilitation", psi = mu_a/2 (sqrt(I5) - S(I4))^2
// with I5 = |F a|^2, I4 = a^T S a (F = R S, R proper), S(x) = -1 if x < 0 else 1:
// P += mu_a (1 - S(I4) / sqrt(I5)) F a a^T, sqrt(I5) clamped to >= 1e-6.  S is
// positive definite when det F > 0 (FP64 det), so only det F <= 0 takes the
// FP64 polar rotation for the sign of I4 = (R a) . (F a).
__device__ __forceinline__ void arap_fiber_add(const float F[3][3], const float a[3], float mu_a, float P[3][3]) {
    float Fa[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) Fa[r] = F[r][0] * a[0] + F[r][1] * a[1] + F[r][2] * a[2];
    double Fd[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Fd[i][j] = (double)F[i][j];
    const double det = Fd[0][0] * (Fd[1][1] * Fd[2][2] - Fd[1][2] * Fd[2][1]) -
                       Fd[0][1] * (Fd[1][0] * Fd[2][2] - Fd[1][2] * Fd[2][0]) +
                       Fd[0][2] * (Fd[1][0] * Fd[2][1] - Fd[1][1] * Fd[2][0]);
    float sgn = 1.0f;
    if (!(det > 0.0)) {
        double Rd[3][3], I4 = 0.0;
        polar_rotation(Fd, Rd);
#pragma unroll
        for (int r = 0; r < 3; ++r)
            I4 += (Rd[r][0] * (double)a[0] + Rd[r][1] * (double)a[1] + Rd[r][2] * (double)a[2]) * (double)Fa[r];
        if (I4 < 0.0) sgn = -1.0f;
    }
    const float r5 = fmaxf(sqrtf(Fa[0] * Fa[0] + Fa[1] * Fa[1] + Fa[2] * Fa[2]), 1e-6f);
    const float k = mu_a * (1.0f - sgn / r5);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) P[r][c] += (k * Fa[r]) * a[c];
}

// Fixed-corotated first Piola-Kirchhoff stress (Stomakhin et al. 2012):
// P = 2 mu (F - R) + lam (J - 1) cof(F), R from the FP64 polar rotation.
__device__ __forceinline__ void fcr_piola(const float F[3][3], float mu, float lam, float P[3][3]) {
    double Fd[3][3], Rd[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Fd[i][j] = (double)F[i][j];
    if (!polar_rotation_newton(Fd, Rd)) polar_rotation(Fd, Rd);
    const float J = F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) -
                    F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0]) +
                    F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]);
    float cof[3][3];
    cof[0][0] = F[1][1] * F[2][2] - F[1][2] * F[2][1];
    cof[0][1] = F[1][2] * F[2][0] - F[1][0] * F[2][2];
    cof[0][2] = F[1][0] * F[2][1] - F[1][1] * F[2][0];
    cof[1][0] = F[0][2] * F[2][1] - F[0][1] * F[2][2];
    cof[1][1] = F[0][0] * F[2][2] - F[0][2] * F[2][0];
    cof[1][2] = F[0][1] * F[2][0] - F[0][0] * F[2][1];
    cof[2][0] = F[0][1] * F[1][2] - F[0][2] * F[1][1];
    cof[2][1] = F[0][2] * F[1][0] - F[0][0] * F[1][2];
    cof[2][2] = F[0][0] * F[1][1] - F[0][1] * F[1][0];
    const float twomu = 2.0f * mu, lj = lam * (J - 1.0f);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) P[i][j] = twomu * (F[i][j] - (float)Rd[i][j]) + lj * cof[i][j];
}

// ---- StVK with Hencky strain and its return mappings
// psi = mu tr(e^2) + lam/2 tr(e)^2 with the Hencky strain e = log(Sigma) (Klar,
// Gast, Pradhana, Fu, Schroeder, Jiang, Teran 2016, "Drucker-Prager
// Elastoplasticity for Sand Animation", sec. 4).  Everything needed comes from
// the left stretch: F F^T = U Sigma^2 U^T (FP64 Jacobi), Sigma clamped to
// kMinStretch as in the reference's Hencky strain of curves
// (S.abs().max(128 eps).log(), MPMPipeline.cu:1037).
constexpr double kMinStretch = 128.0 * 1.1920928955078125e-7;

__device__ __forceinline__ void left_stretch(const float F[3][3], double U[3][3], double e[3], double s[3]) {
    double B[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            B[i][j] = (double)F[i][0] * F[j][0] + (double)F[i][1] * F[j][1] + (double)F[i][2] * F[j][2];
    jacobi_eig3(B, U);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        s[k] = sqrt(fmax(B[k][k], kMinStretch * kMinStretch));
        e[k] = log(s[k]);
    }
}

// Kirchhoff stress tau = P F^T = U diag(2 mu e + lam tr(e)) U^T.
__device__ __forceinline__ void stvk_hencky_tau(const float F[3][3], float mu, float lam, float tau[3][3]) {
    double U[3][3], e[3], s[3];
    left_stretch(F, U, e, s);
    const double tr = e[0] + e[1] + e[2];
    double d[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) d[k] = 2.0 * (double)mu * e[k] + (double)lam * tr;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            tau[i][j] = (float)(U[i][0] * d[0] * U[j][0] + U[i][1] * d[1] * U[j][1] + U[i][2] * d[2] * U[j][2]);
}

// Return mapping of the Hencky strain (ZSReturnMapping, MPMPipeline.cu:953-972):
// the projected strain e' gives F' = U diag(exp(e') / Sigma) U^T F (the right
// factor of F is kept).  Drucker-Prager (Klar et al. 2016, sec. 7.3,
// non-associative): tr(e) >= 0 resets e' = 0; otherwise
// dg = |e^| + (3 lam + 2 mu) / (2 mu) tr(e) alpha and, when dg > 0,
// e' = e - dg e^ / |e^|, with e^ the deviatoric part.  Von Mises (Jiang,
// Schroeder, Teran, Stomakhin, Selle 2016, "The Material Point Method for
// Simulating Continuum Materials", SIGGRAPH course, sec. 9): dg = |e^| -
// yield / (2 mu) and the same e' when dg > 0.  Returns whether F changed.
__device__ __forceinline__ bool project_hencky_strain(float F[3][3], float mu, float lam, int plasticity,
                                                      float dp_alpha, float yield_stress) {
    double U[3][3], e[3], s[3];
    left_stretch(F, U, e, s);
    const double tr = e[0] + e[1] + e[2];
    const double eh[3] = {e[0] - tr / 3.0, e[1] - tr / 3.0, e[2] - tr / 3.0};
    const double nh = sqrt(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]);
    double ep[3];
    if (plasticity == ZMPM_PLASTIC_DRUCKER_PRAGER) {
        if (tr >= 0.0) {
            ep[0] = ep[1] = ep[2] = 0.0;
        } else {
            const double dg = nh + (3.0 * lam + 2.0 * mu) / (2.0 * mu) * tr * (double)dp_alpha;
            if (!(dg > 0.0)) return false;
#pragma unroll
            for (int k = 0; k < 3; ++k) ep[k] = e[k] - dg / nh * eh[k];
        }
    } else if (plasticity == ZMPM_PLASTIC_VON_MISES) {
        const double dg = nh - (double)yield_stress / (2.0 * mu);
        if (!(dg > 0.0)) return false;
#pragma unroll
        for (int k = 0; k < 3; ++k) ep[k] = e[k] - dg / nh * eh[k];
    } else {
        return false;
    }
    double r[3], M[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) r[k] = exp(ep[k]) / s[k];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) M[i][j] = U[i][0] * r[0] * U[j][0] + U[i][1] * r[1] * U[j][1] + U[i][2] * r[2] * U[j][2];
    float Fn[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Fn[i][j] = (float)(M[i][0] * F[0][j] + M[i][1] * F[1][j] + M[i][2] * F[2][j]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) F[i][j] = Fn[i][j];
    return true;
}

// Non-associative Cam-clay on the Hencky strain ("nacc", Generation.cu:77-82;
// zpc's NonAssociativeCamClay is absent, so this restates the published model
// of Gaume, Gast, Teran, van Herwijnen, Jiang 2018, "Dynamic anticrack
// propagation in snow", Nat. Commun. 9:3047, Methods): with the Kirchhoff
// stress tau = 2 mu e + lam tr(e) I of StVK-Hencky, p = -tr(tau)/3 = -K tr(e)
// (K = 2 mu / 3 + lam), s = 2 mu e^ and q^2 = 3/2 |s|^2, the yield surface is
// y = (1 + 2 beta) q^2 + M^2 (p + beta p0)(p - p0) <= 0 with the hardened
// p0 = K (1e-5 + sinh(xi max(-logJp, 0))).  Outside it, a trial pressure above
// p0 or below -beta p0 goes to that tip (isotropic strain, s = 0); otherwise
// p is kept and s scaled onto the surface (non-associative).  logJp gains the
// volume the projection removes, tr(e) - tr(e').  Returns whether F changed.
__device__ __forceinline__ bool project_cam_clay(float F[3][3], float mu, float M, float K, float beta, float xi,
                                                 float &logJp) {
    double U[3][3], e[3], s[3];
    left_stretch(F, U, e, s);
    const double tr = e[0] + e[1] + e[2];
    const double eh[3] = {e[0] - tr / 3.0, e[1] - tr / 3.0, e[2] - tr / 3.0};
    const double nh = sqrt(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]);
    const double Kd = (double)K, Md = (double)M, bd = (double)beta;
    const double p0 = Kd * (1e-5 + sinh((double)xi * fmax(-(double)logJp, 0.0)));
    const double p = -Kd * tr;
    const double q2 = 1.5 * (2.0 * (double)mu * nh) * (2.0 * (double)mu * nh);
    const double y = (1.0 + 2.0 * bd) * q2 + Md * Md * (p + bd * p0) * (p - p0);
    if (!(y > 0.0)) return false;
    double ep[3], trn;
    if (p > p0) {
        trn = -p0 / Kd;
        ep[0] = ep[1] = ep[2] = trn / 3.0;
    } else if (p < -bd * p0) {
        trn = bd * p0 / Kd;
        ep[0] = ep[1] = ep[2] = trn / 3.0;
    } else {
        const double qn2 = fmax(-Md * Md * (p + bd * p0) * (p - p0) / (1.0 + 2.0 * bd), 0.0);
        const double scale = sqrt(qn2 / q2);
        trn = tr;
#pragma unroll
        for (int k = 0; k < 3; ++k) ep[k] = tr / 3.0 + scale * eh[k];
    }
    logJp = (float)((double)logJp + (tr - trn));
    double r[3], Mx[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) r[k] = exp(ep[k]) / s[k];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            Mx[i][j] = U[i][0] * r[0] * U[j][0] + U[i][1] * r[1] * U[j][1] + U[i][2] * r[2] * U[j][2];
    float Fn[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Fn[i][j] = (float)(Mx[i][0] * F[0][j] + Mx[i][1] * F[1][j] + Mx[i][2] * F[2][j]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) F[i][j] = Fn[i][j];
    return true;
}

// Fixed-corotated stress with its pieces kept for the adjoint: P, the FP64
// rotation R, J and cof(F).
__device__ __forceinline__ void fcr_piola_parts(const float F[3][3], float mu, float lam, float P[3][3],
                                                double Rd[3][3], float &J, float cof[3][3]) {
    double Fd[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Fd[i][j] = (double)F[i][j];
    if (!polar_rotation_newton(Fd, Rd)) polar_rotation(Fd, Rd);
    J = F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) - F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0]) +
        F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]);
    cof[0][0] = F[1][1] * F[2][2] - F[1][2] * F[2][1];
    cof[0][1] = F[1][2] * F[2][0] - F[1][0] * F[2][2];
    cof[0][2] = F[1][0] * F[2][1] - F[1][1] * F[2][0];
    cof[1][0] = F[0][2] * F[2][1] - F[0][1] * F[2][2];
    cof[1][1] = F[0][0] * F[2][2] - F[0][2] * F[2][0];
    cof[1][2] = F[0][1] * F[2][0] - F[0][0] * F[2][1];
    cof[2][0] = F[0][1] * F[1][2] - F[0][2] * F[1][1];
    cof[2][1] = F[0][2] * F[1][0] - F[0][0] * F[1][2];
    cof[2][2] = F[0][0] * F[1][1] - F[0][1] * F[1][0];
    const float twomu = 2.0f * mu, lj = lam * (J - 1.0f);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) P[i][j] = twomu * (F[i][j] - (float)Rd[i][j]) + lj * cof[i][j];
}

// Adjoint of P(F) = 2 mu (F - R) + lam (J - 1) cof(F) (FP64 inside): given Pb = dL/dP,
// adds dL/dF to Fb and returns dL/dmu, dL/dlam.  The rotation term uses the
// polar differential: with S = R^T F and M = tr(S) I - S, a variation dF moves R
// by R [w]x with M w = axial(R^T dF - dF^T R), so dL/dF += 2 R [M^-1 a]x with
// a = axial(skew(R^T Rb)) and Rb = -2 mu Pb.  cof(F) has rows r1 x r2, r2 x r0,
// r0 x r1 of the rows r_k of F; J varies by cof(F) : dF.
__device__ __forceinline__ void fcr_piola_adjoint(const float F[3][3], const double R[3][3], float J,
                                                  const float cof[3][3], float mu, float lam, const float Pb[3][3],
                                                  float Fb[3][3], float &mub, float &lamb) {
    double Pbd[3][3], Fd[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            Pbd[i][j] = (double)Pb[i][j];
            Fd[i][j] = (double)F[i][j];
        }
    double acc[3][3];
    double pc = 0.0, pfr = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            acc[i][j] = 2.0 * mu * Pbd[i][j];
            pc += Pbd[i][j] * (double)cof[i][j];
            pfr += Pbd[i][j] * (Fd[i][j] - R[i][j]);
        }
    mub = (float)(2.0 * pfr);
    lamb = (float)(((double)J - 1.0) * pc);
    // rotation: A = R^T Rb with Rb = -2 mu Pb
    double A[3][3], S[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            A[i][j] = -2.0 * mu * (R[0][i] * Pbd[0][j] + R[1][i] * Pbd[1][j] + R[2][i] * Pbd[2][j]);
            S[i][j] = R[0][i] * Fd[0][j] + R[1][i] * Fd[1][j] + R[2][i] * Fd[2][j];
        }
    const double a[3] = {0.5 * (A[2][1] - A[1][2]), 0.5 * (A[0][2] - A[2][0]), 0.5 * (A[1][0] - A[0][1])};
    const double tr = S[0][0] + S[1][1] + S[2][2];
    double M[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) M[i][j] = (i == j ? tr : 0.0) - 0.5 * (S[i][j] + S[j][i]);
    // c = M^-1 a (M is symmetric positive definite for det F > 0; guarded otherwise)
    const double m00 = M[1][1] * M[2][2] - M[1][2] * M[2][1], m01 = M[0][2] * M[2][1] - M[0][1] * M[2][2],
                 m02 = M[0][1] * M[1][2] - M[0][2] * M[1][1], m11 = M[0][0] * M[2][2] - M[0][2] * M[2][0],
                 m12 = M[0][2] * M[1][0] - M[0][0] * M[1][2], m22 = M[0][0] * M[1][1] - M[0][1] * M[1][0];
    const double det = M[0][0] * m00 + M[0][1] * (M[1][2] * M[2][0] - M[1][0] * M[2][2]) + M[0][2] *
                       (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    double c[3] = {0.0, 0.0, 0.0};
    if (fabs(det) > 1e-30) {
        const double inv = 1.0 / det;
        c[0] = (m00 * a[0] + m01 * a[1] + m02 * a[2]) * inv;
        c[1] = (m01 * a[0] + m11 * a[1] + m12 * a[2]) * inv;
        c[2] = (m02 * a[0] + m12 * a[1] + m22 * a[2]) * inv;
    }
    const double Ch[3][3] = {{0.0, -c[2], c[1]}, {c[2], 0.0, -c[0]}, {-c[1], c[0], 0.0}};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) acc[i][j] += 2.0 * (R[i][0] * Ch[0][j] + R[i][1] * Ch[1][j] + R[i][2] * Ch[2][j]);
    // lam (J - 1) cof(F): through J, then through cof
    const double lpc = lam * pc, lj = lam * ((double)J - 1.0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) acc[i][j] += lpc * (double)cof[i][j];
    // row k of cof = r_{k+1} x r_{k+2}; its adjoint row q adds r_{k+2} x q to r_{k+1} and q x r_{k+1} to r_{k+2}
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
        const double q[3] = {lj * Pbd[k][0], lj * Pbd[k][1], lj * Pbd[k][2]};
        const double *r1 = Fd[k1], *r2 = Fd[k2];
        acc[k1][0] += r2[1] * q[2] - r2[2] * q[1];
        acc[k1][1] += r2[2] * q[0] - r2[0] * q[2];
        acc[k1][2] += r2[0] * q[1] - r2[1] * q[0];
        acc[k2][0] += q[1] * r1[2] - q[2] * r1[1];
        acc[k2][1] += q[2] * r1[0] - q[0] * r1[2];
        acc[k2][2] += q[0] * r1[1] - q[1] * r1[0];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Fb[i][j] += (float)acc[i][j];
}

}  // namespace zmpm
