/*
 * a2b_oracle.c -- FP64 CPU oracle for ReducedModel.Angles2Backbone.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (torchproteinlibrary_amd/)
 * links, loads or calls this file.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, as the checker and as the reported CPU
 * baseline.
 *
 * It restates the reference's CPU algorithm from the math, in plain C:
 *   - forward:  Layers/ReducedModel/cBackboneProteinCPUKernels.cpp:14-60
 *               (A_0 = I, x_0 = 0, A_i = A_{i-1} B_i, x_i = A_i 0bar; the
 *               angle->atom map of :42-51; constants of :6-12; the B(theta,
 *               beta, d) element layout of Math/cMatrix44.cpp:47-52)
 *   - backward, per-pair form: cBackboneProteinCPUKernels.cpp:62-145
 *               (dx_i/dtheta = A_{m-1} dB_m A_m^{-1} A_i 0bar, the dB layout of
 *               Math/cMatrix44.cpp:54-59, the rigid inverse of
 *               Math/cMatrix44.cpp:142-153), with the out-of-chain reads of the
 *               last psi/omega replaced by their true value 0.
 *   - backward, dense form: the same per-pair values materialised into a
 *               zero-filled [B,3,9*L*L] tensor and then reduced, i.e. the memory
 *               behaviour of the reference CPU path
 *               (TorchProteinLibrary/ReducedModel/Angles2Backbone/Angles2Backbone.py:80-88).
 *   - backward, O(L) adjoint: grad_m = u_m . sum_{i>m} (x_i - x_m) x g_i, the
 *               closed form of the same derivative (docs/reducedmodel.html:94-166
 *               gives dx_i/dtheta; the rotation-axis form follows from it).
 *
 * Layouts follow the reference API: angles [B][3][L] (phi, psi, omega rows),
 * lengths int32 [B], coords [B][9L] (atoms N0,CA0,C0,N1,... xyz interleaved),
 * grad_angles [B][3][L].  Every output slot outside a chain is written 0.
 * Lengths are clamped to [0, L].
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define A2BO_R_N_CA 1.460
#define A2BO_R_CA_C 1.525
#define A2BO_R_C_N 1.330
/* beta = pi - bond angle (cBackboneProteinCPUKernels.cpp:10-12); M_PI value */
#define A2BO_PI 3.14159265358979323846
#define A2BO_C_N_CA (A2BO_PI - 1.9391)
#define A2BO_N_CA_C (A2BO_PI - 2.061)
#define A2BO_CA_C_N (A2BO_PI - 2.1186)

/* 4x4 row-major homogeneous matrices, as in Math/cMatrix44.h. */
typedef struct { double m[16]; } mat4;

static void m4_identity(mat4 *a) {
    memset(a->m, 0, sizeof(a->m));
    a->m[0] = a->m[5] = a->m[10] = a->m[15] = 1.0;
}

static void m4_mul(const mat4 *a, const mat4 *b, mat4 *c) {
    mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += a->m[4 * i + k] * b->m[4 * k + j];
            r.m[4 * i + j] = s;
        }
    *c = r;
}

/* B(theta, beta, d) = Ry(beta) Tx(d) Rx(theta): Math/cMatrix44.cpp:47-52. */
static void m4_dihedral(mat4 *b, double th, double be, double d) {
    double ct = cos(th), st = sin(th), cb = cos(be), sb = sin(be);
    double *m = b->m;
    m[0] = cb;  m[1] = st * sb; m[2] = ct * sb;  m[3] = d * cb;
    m[4] = 0.0; m[5] = ct;      m[6] = -st;      m[7] = 0.0;
    m[8] = -sb; m[9] = st * cb; m[10] = ct * cb; m[11] = -d * sb;
    m[12] = 0.0; m[13] = 0.0; m[14] = 0.0; m[15] = 1.0;
}

/* dB/dtheta: Math/cMatrix44.cpp:54-59. */
static void m4_dihedral_dtheta(mat4 *b, double th, double be) {
    double ct = cos(th), st = sin(th), cb = cos(be), sb = sin(be);
    double *m = b->m;
    memset(m, 0, sizeof(b->m));
    m[1] = ct * sb; m[2] = -st * sb;
    m[5] = -st;     m[6] = -ct;
    m[9] = ct * cb; m[10] = -st * cb;
}

/* Rigid inverse [R^T, -R^T t]: Math/cMatrix44.cpp:142-153. */
static void m4_rigid_inverse(const mat4 *a, mat4 *r) {
    mat4 o;
    memset(o.m, 0, sizeof(o.m));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) o.m[4 * i + j] = a->m[4 * j + i];
    for (int i = 0; i < 3; ++i)
        o.m[4 * i + 3] = -(o.m[4 * i + 0] * a->m[3] + o.m[4 * i + 1] * a->m[7] +
                           o.m[4 * i + 2] * a->m[11]);
    o.m[15] = 1.0;
    *r = o;
}

static int clamp_len(int32_t n, int L) { return n < 0 ? 0 : (n > L ? L : n); }

/* Per-atom bond parameters: atom i>=1 uses angle row k=(i-1)%3 of residue
 * (i-1)/3 (cBackboneProteinCPUKernels.cpp:42-51: i%3==1 -> phi[i/3],
 * i%3==2 -> psi[i/3], i%3==0 -> omega[i/3-1]). */
static void atom_params(int i, int *k, int *j, double *be, double *d) {
    *k = (i - 1) % 3;
    *j = (i - 1) / 3;
    static const double bes[3] = {A2BO_C_N_CA, A2BO_N_CA_C, A2BO_CA_C_N};
    static const double ds[3] = {A2BO_R_N_CA, A2BO_R_CA_C, A2BO_R_C_N};
    *be = bes[*k];
    *d = ds[*k];
}

/* Frames of one chain: A[i] for i in [0, 3n).  Returns the frames array
 * (caller frees). cBackboneProteinCPUKernels.cpp:36-58. */
static mat4 *chain_frames(const double *ang, int L, int n) {
    int na = 3 * n;
    mat4 *A = (mat4 *)malloc(sizeof(mat4) * (size_t)(na > 0 ? na : 1));
    if (na == 0) return A;
    m4_identity(&A[0]);
    for (int i = 1; i < na; ++i) {
        int k, j;
        double be, d;
        atom_params(i, &k, &j, &be, &d);
        mat4 B;
        m4_dihedral(&B, ang[k * L + j], be, d);
        m4_mul(&A[i - 1], &B, &A[i]);
    }
    return A;
}

void a2bo_forward(const double *angles, const int32_t *lengths, double *coords,
                  int B, int L) {
    for (int b = 0; b < B; ++b) {
        const double *ang = angles + (size_t)b * 3 * L;
        double *x = coords + (size_t)b * 9 * L;
        memset(x, 0, sizeof(double) * 9 * (size_t)L);
        int n = clamp_len(lengths[b], L);
        mat4 *A = chain_frames(ang, L, n);
        for (int i = 0; i < 3 * n; ++i) {
            x[3 * i + 0] = A[i].m[3];
            x[3 * i + 1] = A[i].m[7];
            x[3 * i + 2] = A[i].m[11];
        }
        free(A);
    }
}

/* dx_i/dtheta_m for one (angle, atom) pair, the reference per-pair formula
 * (cBackboneProteinCPUKernels.cpp:78-90): A_{m-1} dB_m A_m^{-1} A_i 0bar,
 * with M = A_{m-1} dB_m A_m^{-1} hoisted by the caller. */
static void pair_apply(const mat4 *M, const mat4 *Ai, double out[3]) {
    for (int r = 0; r < 3; ++r)
        out[r] = M->m[4 * r + 0] * Ai->m[3] + M->m[4 * r + 1] * Ai->m[7] +
                 M->m[4 * r + 2] * Ai->m[11] + M->m[4 * r + 3];
}

static void hoisted_M(const mat4 *A, const double *ang, int L, int m, mat4 *M) {
    int k, j;
    double be, d;
    atom_params(m, &k, &j, &be, &d);
    mat4 dB, inv, tmp;
    m4_dihedral_dtheta(&dB, ang[k * L + j], be);
    m4_rigid_inverse(&A[m], &inv);
    m4_mul(&A[m - 1], &dB, &tmp);
    m4_mul(&tmp, &inv, M);
}

/* O(L^2) per-pair backward, summed directly (no derivative tensor). */
void a2bo_backward_pairwise(const double *angles, const int32_t *lengths,
                            const double *grad_coords, double *grad_angles,
                            int B, int L) {
    for (int b = 0; b < B; ++b) {
        const double *ang = angles + (size_t)b * 3 * L;
        const double *g = grad_coords + (size_t)b * 9 * L;
        double *ga = grad_angles + (size_t)b * 3 * L;
        memset(ga, 0, sizeof(double) * 3 * (size_t)L);
        int n = clamp_len(lengths[b], L);
        int na = 3 * n;
        mat4 *A = chain_frames(ang, L, n);
        /* angle (k, j) moves atoms i > m = 3j+k+1; m >= na-1 has none. */
        for (int m = 1; m < na - 1; ++m) {
            mat4 M;
            hoisted_M(A, ang, L, m, &M);
            double s = 0.0;
            for (int i = m + 1; i < na; ++i) {
                double dx[3];
                pair_apply(&M, &A[i], dx);
                s += g[3 * i] * dx[0] + g[3 * i + 1] * dx[1] + g[3 * i + 2] * dx[2];
            }
            int k = (m - 1) % 3, j = (m - 1) / 3;
            ga[k * L + j] = s;
        }
        free(A);
    }
}

/* The reference CPU algorithm's memory behaviour: materialise every per-pair
 * derivative into a zero-filled [B][3][9 L L] tensor
 * (cBackboneProteinCPUKernels.cpp:93-117 with the index map of :110-111), then
 * reduce it (:119-145).  dr_dangle must hold B*3*9*L*L doubles; it is
 * zero-filled here, as Angles2Backbone.py:84 does. */
void a2bo_backward_dense(const double *angles, const int32_t *lengths,
                         const double *grad_coords, double *grad_angles,
                         double *dr_dangle, int B, int L) {
    size_t per_k = (size_t)9 * L * L; /* atoms_stride*angles_stride*3 */
    memset(dr_dangle, 0, sizeof(double) * (size_t)B * 3 * per_k);
    for (int b = 0; b < B; ++b) {
        const double *ang = angles + (size_t)b * 3 * L;
        int n = clamp_len(lengths[b], L);
        int na = 3 * n;
        mat4 *A = chain_frames(ang, L, n);
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < 3; ++k) {
                int m = 3 * j + k + 1;
                if (m >= na - 1) continue; /* true derivative is 0 */
                mat4 M;
                hoisted_M(A, ang, L, m, &M);
                double *dst = dr_dangle + ((size_t)3 * b + k) * per_k + (size_t)j * 9 * L;
                for (int i = m + 1; i < na; ++i) pair_apply(&M, &A[i], dst + 3 * i);
            }
        free(A);
    }
    for (int b = 0; b < B; ++b) {
        int n = clamp_len(lengths[b], L);
        const double *g = grad_coords + (size_t)b * 9 * L;
        double *ga = grad_angles + (size_t)b * 3 * L;
        memset(ga, 0, sizeof(double) * 3 * (size_t)L);
        for (int k = 0; k < 3; ++k)
            for (int j = 0; j < n; ++j) {
                const double *src = dr_dangle + ((size_t)3 * b + k) * per_k + (size_t)j * 9 * L;
                double s = 0.0;
                for (int i = 3 * j; i < 3 * n; ++i)
                    s += g[3 * i] * src[3 * i] + g[3 * i + 1] * src[3 * i + 1] +
                         g[3 * i + 2] * src[3 * i + 2];
                ga[k * L + j] = s;
            }
    }
}

/* O(L) adjoint: grad_m = u_m . (T_m - x_m x F_m) with F_m = sum_{i>m} g_i,
 * T_m = sum_{i>m} x_i x g_i and u_m the unit x-axis of frame m (the bond
 * direction x_m - x_{m-1}).  Sums are accumulated in long double. */
void a2bo_backward_adjoint(const double *angles, const int32_t *lengths,
                           const double *grad_coords, double *grad_angles,
                           int B, int L) {
    for (int b = 0; b < B; ++b) {
        const double *ang = angles + (size_t)b * 3 * L;
        const double *g = grad_coords + (size_t)b * 9 * L;
        double *ga = grad_angles + (size_t)b * 3 * L;
        memset(ga, 0, sizeof(double) * 3 * (size_t)L);
        int n = clamp_len(lengths[b], L);
        int na = 3 * n;
        mat4 *A = chain_frames(ang, L, n);
        long double F[3] = {0, 0, 0}, T[3] = {0, 0, 0};
        for (int m = na - 1; m >= 1; --m) {
            if (m < na - 1) {
                int i = m + 1;
                long double xi[3] = {A[i].m[3], A[i].m[7], A[i].m[11]};
                long double gi[3] = {g[3 * i], g[3 * i + 1], g[3 * i + 2]};
                F[0] += gi[0]; F[1] += gi[1]; F[2] += gi[2];
                T[0] += xi[1] * gi[2] - xi[2] * gi[1];
                T[1] += xi[2] * gi[0] - xi[0] * gi[2];
                T[2] += xi[0] * gi[1] - xi[1] * gi[0];
                long double xm[3] = {A[m].m[3], A[m].m[7], A[m].m[11]};
                long double u[3] = {A[m].m[0], A[m].m[4], A[m].m[8]};
                long double w[3] = {T[0] - (xm[1] * F[2] - xm[2] * F[1]),
                                    T[1] - (xm[2] * F[0] - xm[0] * F[2]),
                                    T[2] - (xm[0] * F[1] - xm[1] * F[0])};
                int k = (m - 1) % 3, j = (m - 1) / 3;
                ga[k * L + j] = (double)(u[0] * w[0] + u[1] * w[1] + u[2] * w[2]);
            }
        }
        free(A);
    }
}

/* Per-atom frames (3x4 row-major: R row 0 | t0, R row 1 | t1, R row 2 | t2)
 * for inspection in tests: out [B][3L][12], zero outside chains. */
void a2bo_frames(const double *angles, const int32_t *lengths, double *out,
                 int B, int L) {
    for (int b = 0; b < B; ++b) {
        const double *ang = angles + (size_t)b * 3 * L;
        double *o = out + (size_t)b * 3 * L * 12;
        memset(o, 0, sizeof(double) * 36 * (size_t)L);
        int n = clamp_len(lengths[b], L);
        mat4 *A = chain_frames(ang, L, n);
        for (int i = 0; i < 3 * n; ++i) memcpy(o + 12 * i, A[i].m, sizeof(double) * 12);
        free(A);
    }
}
