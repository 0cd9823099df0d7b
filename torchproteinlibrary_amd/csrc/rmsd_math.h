// rmsd_math.h -- the per-structure eigen-solve of Coords2RMSD, in FP64.  Host
// and device code: the HIP kernels use it, and a CPU test compiles the same
// header with g++ and checks it against numpy.
//
// For the centred correlation R (3x3, R[r][c] = sum x_r y_c) the reference
// builds the symmetric 4x4 matrix F (coords2rmsd_interface.cpp:155-158); its
// largest eigenvalue lambda gives rmsd^2 = (R2 - 2 lambda) / n and its unit
// eigenvector q the rotation U(q) (:35-45).  The reference takes both from a
// general eigen-solver (:12-32).  Here:
//   * lambda by Newton's method on the characteristic quartic of F, started
//     from R2 / 2, an upper bound of lambda (the "QCP" method: D. L. Theobald,
//     Acta Cryst. A61 (2005) 478-480).  F is symmetric, so every root is
//     real and Newton from above the largest root decreases monotonically
//     onto it;
//   * q from a column of adj(F - mu I), the column of the diagonal entry
//     largest in magnitude.  For a simple eigenvalue mu = lambda, where
//     adj(F - lambda I) = prod_{j != max} (lambda_j - lambda) q q^T
//     = -P'(lambda) q q^T: the three factors are negative, so every diagonal
//     entry -P'(lambda) q_i^2 is <= 0 and the most negative one belongs to
//     the largest |q_i|.  When the largest eigenvalue is (nearly) double,
//     P'(lambda) vanishes and that adjugate carries no direction; then mu =
//     lambda (1 + 1e-8), where adj(F - mu I) = sum_k prod_{j!=k} (lambda_j -
//     mu) q_k q_k^T is dominated by the top eigenspace (its terms are O(delta)
//     against O(delta^2) for the others): any of its columns is an optimal
//     quaternion, up to O(1e-8) of the other eigenvectors.  There lambda
//     itself is refined by Newton on P', whose root is simple.
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define RMSD_HD __host__ __device__ __forceinline__
#define RMSD_UNROLL _Pragma("unroll")
#else
#define RMSD_HD inline
#define RMSD_UNROLL
#endif

namespace rmsd_math {

// Largest eigenvalue, its unit eigenvector (a quaternion), and whether it was
// (nearly) double.
struct Eigenpair {
    double lam, q0, q1, q2, q3;
    int degenerate;
};

// F from R (both row-major), in the reference's element layout.
RMSD_HD void f_matrix(const double R[9], double F[16]) {
    F[0] = R[0] + R[4] + R[8];
    F[1] = R[5] - R[7];
    F[2] = R[6] - R[2];
    F[3] = R[1] - R[3];
    F[4] = F[1];
    F[5] = R[0] - R[4] - R[8];
    F[6] = R[1] + R[3];
    F[7] = R[2] + R[6];
    F[8] = F[2];
    F[9] = F[6];
    F[10] = -R[0] + R[4] - R[8];
    F[11] = R[5] + R[7];
    F[12] = F[3];
    F[13] = F[7];
    F[14] = F[11];
    F[15] = -R[0] - R[4] + R[8];
}

// Rotation matrix (row-major) of the unit quaternion (q0, q1, q2, q3) (the
// reference's formula).
RMSD_HD void u_from_q(const Eigenpair& e, double U[9]) {
    const double q0 = e.q0, q1 = e.q1, q2 = e.q2, q3 = e.q3;
    U[0] = q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3;
    U[1] = 2.0 * (q1 * q2 - q0 * q3);
    U[2] = 2.0 * (q1 * q3 + q0 * q2);
    U[3] = 2.0 * (q1 * q2 + q0 * q3);
    U[4] = q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3;
    U[5] = 2.0 * (q2 * q3 - q0 * q1);
    U[6] = 2.0 * (q1 * q3 - q0 * q2);
    U[7] = 2.0 * (q2 * q3 + q0 * q1);
    U[8] = q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3;
}

RMSD_HD double det3(double a, double b, double c, double d, double e, double f, double g, double h,
                    double i) {
    return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
}

// Minor of the 4x4 matrix A without row R and column K (template indices,
// so every element index is a constant).
template <int R, int K>
RMSD_HD double minor3(const double A[16]) {
    constexpr int r0 = R == 0 ? 1 : 0, r1 = R <= 1 ? 2 : 1, r2 = R <= 2 ? 3 : 2;
    constexpr int c0 = K == 0 ? 1 : 0, c1 = K <= 1 ? 2 : 1, c2 = K <= 2 ? 3 : 2;
    return det3(A[4 * r0 + c0], A[4 * r0 + c1], A[4 * r0 + c2], A[4 * r1 + c0], A[4 * r1 + c1],
                A[4 * r1 + c2], A[4 * r2 + c0], A[4 * r2 + c1], A[4 * r2 + c2]);
}

// Row R of the adjugate of the symmetric A (= its column R): the cofactors
// (-1)^(R+K) minor(R, K), into four scalars.
template <int R>
RMSD_HD void cofactor_row(const double A[16], double& c0, double& c1, double& c2, double& c3) {
    c0 = ((R + 0) & 1) ? -minor3<R, 0>(A) : minor3<R, 0>(A);
    c1 = ((R + 1) & 1) ? -minor3<R, 1>(A) : minor3<R, 1>(A);
    c2 = ((R + 2) & 1) ? -minor3<R, 2>(A) : minor3<R, 2>(A);
    c3 = ((R + 3) & 1) ? -minor3<R, 3>(A) : minor3<R, 3>(A);
}

// Largest eigenvalue of F and a unit eigenvector.  `bound` >= the largest
// eigenvalue (R2 / 2 for the F of centred structures).
RMSD_HD Eigenpair max_eigenpair(const double F[16], double bound) {
    // det(lambda I - F) = lambda^4 + c2 lambda^2 + c1 lambda + c0 (trace F = 0):
    // c2 = sum of the principal 2x2 minors, c1 = -(sum of the principal 3x3
    // minors), c0 = det F.
    const double c2 = F[0] * F[5] + F[0] * F[10] + F[0] * F[15] + F[5] * F[10] + F[5] * F[15] +
                      F[10] * F[15] - F[1] * F[1] - F[2] * F[2] - F[3] * F[3] - F[6] * F[6] -
                      F[7] * F[7] - F[11] * F[11];
    const double m00 = minor3<0, 0>(F), m01 = minor3<0, 1>(F), m02 = minor3<0, 2>(F),
                 m03 = minor3<0, 3>(F);
    const double c1 = -(m00 + minor3<1, 1>(F) + minor3<2, 2>(F) + minor3<3, 3>(F));
    const double c0 = F[0] * m00 - F[1] * m01 + F[2] * m02 - F[3] * m03;
    double lam = bound;
    if (!(lam > 0.0)) return Eigenpair{0.0, 1.0, 0.0, 0.0, 0.0, 0};  // F = 0: the identity
    for (int it = 0; it < 60; ++it) {
        const double p = ((lam * lam + c2) * lam + c1) * lam + c0;
        const double dp = (4.0 * lam * lam + 2.0 * c2) * lam + c1;
        if (!(dp > 0.0)) break;
        const double d = p / dp;
        lam -= d;
        if (fabs(d) <= 1e-14 * fabs(lam)) break;
    }
    // P'(lambda) = prod over the other eigenvalues of (lambda - lambda_j);
    // the other eigenvalues lie in [-3 lambda, lambda] (trace 0).
    const double dp = (4.0 * lam * lam + 2.0 * c2) * lam + c1;
    const int degenerate = !(dp > 1e-6 * lam * lam * lam);
    if (degenerate) {
        // Near a double root P falls below its rounding noise and Newton on P
        // stalls near sqrt(eps); the root of P' there is simple: Newton on P'.
        for (int it = 0; it < 8; ++it) {
            const double d1 = (4.0 * lam * lam + 2.0 * c2) * lam + c1;
            const double d2 = 12.0 * lam * lam + 2.0 * c2;
            if (!(d2 > 0.0)) break;
            const double d = d1 / d2;
            lam -= d;
            if (fabs(d) <= 1e-15 * fabs(lam)) break;
        }
    }
    const double mu = degenerate ? lam * (1.0 + 1e-8) : lam;
    double A[16];
    RMSD_UNROLL
    for (int i = 0; i < 16; ++i) A[i] = F[i] - ((i % 5 == 0) ? mu : 0.0);
    // adj(A) row i has diagonal entry minor(i, i) = -P'(lambda) q_i^2 <= 0
    // (mu >= every eigenvalue, so each of the three factors lambda_j - mu is
    // <= 0, in both branches): the most negative row carries q best; a row
    // with q_i = 0 holds only rounding noise.
    const double d0 = minor3<0, 0>(A), d1 = minor3<1, 1>(A), d2 = minor3<2, 2>(A),
                 d3 = minor3<3, 3>(A);
    int r = 0;
    double best = d0;
    if (d1 < best) { best = d1; r = 1; }
    if (d2 < best) { best = d2; r = 2; }
    if (d3 < best) r = 3;
    double w0, w1, w2, w3;
    switch (r) {
        case 0: cofactor_row<0>(A, w0, w1, w2, w3); break;
        case 1: cofactor_row<1>(A, w0, w1, w2, w3); break;
        case 2: cofactor_row<2>(A, w0, w1, w2, w3); break;
        default: cofactor_row<3>(A, w0, w1, w2, w3); break;
    }
    const double inv = 1.0 / sqrt(w0 * w0 + w1 * w1 + w2 * w2 + w3 * w3);
    return Eigenpair{lam, w0 * inv, w1 * inv, w2 * inv, w3 * inv, degenerate};
}

}  // namespace rmsd_math
