// Coords2RMSD for MI355X (gfx950): minimal RMSD over rigid rotations and its
// gradient, each direction one kernel launch.  The C ABI is include/tpl_rmsd.h.
//
// The layer follows the reference's method (quaternion eigenproblem of the
// centred correlation matrix; coords2rmsd_interface.cpp:12-49,52-111,
// RMSDKernels.cu:5-39) with everything on the device:
//   forward, one workgroup per structure:
//     1. FP64 moments of the atoms relative to the structure's first atom
//        (sum x, sum y, sum x y^T, sum |x|^2, sum |y|^2), summed over lanes
//        (DPP-free lane butterflies through ds_bpermute) and waves (LDS);
//     2. lane 0 centres them, builds the symmetric 4x4 matrix F of the
//        centred correlation R (the reference's element layout,
//        coords2rmsd_interface.cpp:153-158), takes its largest eigenpair by
//        cyclic Jacobi rotations, forms U from the quaternion (same formula,
//        :35-45) and rmsd = sqrt((R2 - 2 |lambda| + 1e-10) / n) (:182);
//   backward, one thread per atom: (x - cx - U^T (y - cy)) * g / (rmsd * n)
//     (Coords2RMSD.py:47-56), zero past num_atoms.
// Shifting by the first atom keeps the raw moments at the size of the
// structure, so centring them afterwards loses no digits to a large offset
// from the origin.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "tpl_rmsd.h"

namespace {

constexpr double kEpsLimit = 1e-10;  // coords2rmsd_interface.cpp:10
constexpr int kThreads = 256;        // forward workgroup: 4 waves
constexpr int kBwdThreads = 256;     // backward: one thread per atom
constexpr int kState = 16;           // doubles per structure: cx[3] cy[3] U[9] rmsd
constexpr int kMoments = 17;         // sx[3] sy[3] sxy[9] sxx syy

__device__ __forceinline__ double dbl_from(int lo, int hi) {
    return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of the wave; every lane gets the total.
__device__ __forceinline__ double wave_sum(double v, int lane) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const int src = (lane ^ m) << 2;
        const int lo = __builtin_amdgcn_ds_bpermute(src, __double2loint(v));
        const int hi = __builtin_amdgcn_ds_bpermute(src, __double2hiint(v));
        v += dbl_from(lo, hi);
    }
    return v;
}

__device__ __forceinline__ void report(uint32_t* status, uint32_t bit) {
    if (status) __hip_atomic_fetch_or(status, bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int clamp_count(int n, int A, uint32_t* status, bool report_it) {
    if (n < 0 || n > A) {
        if (report_it) report(status, TPL_RMSD_STATUS_COUNT);
        return n < 0 ? 0 : A;
    }
    return n;
}

// One Jacobi rotation of the symmetric 4x4 matrix a (row-major) in the plane
// (p, q), chosen to zero a[p][q]; v accumulates the rotations (columns are
// the eigenvectors).  With theta = (a_qq - a_pp) / (2 a_pq), t = tan(phi) is
// the smaller root of t^2 + 2 theta t - 1 = 0.
template <int P, int Q>
__device__ __forceinline__ void jacobi_rotate(double (&a)[16], double (&v)[16]) {
    const double apq = a[4 * P + Q];
    if (apq == 0.0) return;
    const double theta = (a[4 * Q + Q] - a[4 * P + P]) / (2.0 * apq);
    const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
    const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
    for (int k = 0; k < 4; ++k) {  // columns p, q
        const double akp = a[4 * k + P], akq = a[4 * k + Q];
        a[4 * k + P] = c * akp - s * akq;
        a[4 * k + Q] = s * akp + c * akq;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {  // rows p, q
        const double apk = a[4 * P + k], aqk = a[4 * Q + k];
        a[4 * P + k] = c * apk - s * aqk;
        a[4 * Q + k] = s * apk + c * aqk;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double vkp = v[4 * k + P], vkq = v[4 * k + Q];
        v[4 * k + P] = c * vkp - s * vkq;
        v[4 * k + Q] = s * vkp + c * vkq;
    }
}

// Largest eigenvalue of the symmetric 4x4 matrix f and a unit eigenvector.
// Cyclic Jacobi sweeps until the off-diagonal mass is below 1e-32 of the
// diagonal's (or vanishes), at most 20 sweeps.
__device__ double max_eigenpair(const double (&f)[16], double (&q)[4]) {
    double a[16], v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        a[i] = f[i];
        v[i] = (i % 5 == 0) ? 1.0 : 0.0;
    }
    for (int sweep = 0; sweep < 20; ++sweep) {
        const double off = a[1] * a[1] + a[2] * a[2] + a[3] * a[3] + a[6] * a[6] + a[7] * a[7] +
                           a[11] * a[11];
        const double dg = a[0] * a[0] + a[5] * a[5] + a[10] * a[10] + a[15] * a[15];
        if (off == 0.0 || off <= 1e-32 * dg) break;
        jacobi_rotate<0, 1>(a, v);
        jacobi_rotate<0, 2>(a, v);
        jacobi_rotate<0, 3>(a, v);
        jacobi_rotate<1, 2>(a, v);
        jacobi_rotate<1, 3>(a, v);
        jacobi_rotate<2, 3>(a, v);
    }
    int best = 0;
    double lam = a[0];
#pragma unroll
    for (int k = 1; k < 4; ++k) {
        if (a[5 * k] > lam) {
            lam = a[5 * k];
            best = k;
        }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        double e = v[4 * j];
#pragma unroll
        for (int k = 1; k < 4; ++k) e = (best == k) ? v[4 * j + k] : e;
        q[j] = e;
    }
    return lam;
}

template <typename T>
__global__ void __launch_bounds__(kThreads) rmsd_fwd(const T* __restrict__ x, const T* __restrict__ y,
                                                     const int32_t* __restrict__ num_atoms,
                                                     T* rmsd_out, T* ut_out, double* state,
                                                     uint32_t* status, int A) {
    __shared__ double s_red[kThreads / 64][kMoments];
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n = clamp_count(num_atoms[b], A, status, tid == 0);
    const T* xb = x + (size_t)b * 3 * A;
    const T* yb = y + (size_t)b * 3 * A;
    double x0[3] = {0.0, 0.0, 0.0}, y0[3] = {0.0, 0.0, 0.0};
    if (n > 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            x0[c] = (double)xb[c];
            y0[c] = (double)yb[c];
        }
    }
    double m[kMoments];
#pragma unroll
    for (int e = 0; e < kMoments; ++e) m[e] = 0.0;
    for (int i = tid; i < n; i += kThreads) {
        double dx[3], dy[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            dx[c] = (double)xb[3 * i + c] - x0[c];
            dy[c] = (double)yb[3 * i + c] - y0[c];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            m[c] += dx[c];
            m[3 + c] += dy[c];
            m[15] += dx[c] * dx[c];
            m[16] += dy[c] * dy[c];
        }
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) m[6 + 3 * r + c] += dx[r] * dy[c];
    }
#pragma unroll
    for (int e = 0; e < kMoments; ++e) m[e] = wave_sum(m[e], lane);
    if (lane == 0) {
#pragma unroll
        for (int e = 0; e < kMoments; ++e) s_red[wave][e] = m[e];
    }
    __syncthreads();
    if (tid != 0) return;

    const int nw = blockDim.x >> 6;
    for (int w = 1; w < nw; ++w) {
#pragma unroll
        for (int e = 0; e < kMoments; ++e) m[e] += s_red[w][e];
    }
    const double nd = (double)n;
    double mx[3], my[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        mx[c] = n > 0 ? m[c] / nd : 0.0;
        my[c] = n > 0 ? m[3 + c] / nd : 0.0;
    }
    // Centred correlation R[r][c] = sum (x - cx)_r (y - cy)_c, and R2.
    double R[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) R[3 * r + c] = m[6 + 3 * r + c] - nd * mx[r] * my[c];
    const double r2 = (m[15] - nd * (mx[0] * mx[0] + mx[1] * mx[1] + mx[2] * mx[2])) +
                      (m[16] - nd * (my[0] * my[0] + my[1] * my[1] + my[2] * my[2]));
    // F: the reference's element layout (coords2rmsd_interface.cpp:155-158).
    const double F[16] = {
        R[0] + R[4] + R[8], R[5] - R[7],        R[6] - R[2],         R[1] - R[3],
        R[5] - R[7],        R[0] - R[4] - R[8], R[1] + R[3],         R[2] + R[6],
        R[6] - R[2],        R[1] + R[3],        -R[0] + R[4] - R[8], R[5] + R[7],
        R[1] - R[3],        R[2] + R[6],        R[5] + R[7],         -R[0] - R[4] + R[8]};
    double q[4];
    const double lam = max_eigenpair(F, q);
    const double U[9] = {
        q[0] * q[0] + q[1] * q[1] - q[2] * q[2] - q[3] * q[3], 2.0 * (q[1] * q[2] - q[0] * q[3]),
        2.0 * (q[1] * q[3] + q[0] * q[2]),
        2.0 * (q[1] * q[2] + q[0] * q[3]), q[0] * q[0] - q[1] * q[1] + q[2] * q[2] - q[3] * q[3],
        2.0 * (q[2] * q[3] - q[0] * q[1]),
        2.0 * (q[1] * q[3] - q[0] * q[2]), 2.0 * (q[2] * q[3] + q[0] * q[1]),
        q[0] * q[0] - q[1] * q[1] - q[2] * q[2] + q[3] * q[3]};
    const double rmsd = sqrt((r2 - 2.0 * fabs(lam) + kEpsLimit) / nd);
    rmsd_out[b] = (T)rmsd;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) ut_out[9 * b + 3 * j + k] = (T)U[3 * k + j];
    if (state) {
        double* s = state + (size_t)b * kState;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            s[c] = x0[c] + mx[c];
            s[3 + c] = y0[c] + my[c];
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) s[6 + k] = U[k];
        s[15] = rmsd;
    }
    if (rmsd != rmsd) report(status, TPL_RMSD_STATUS_NAN);
}

// One thread per atom slot; the grid covers nblk blocks per structure.
template <typename T>
__global__ void __launch_bounds__(kBwdThreads) rmsd_bwd(T* gx, const T* __restrict__ grad_rmsd,
                                                        const T* __restrict__ x,
                                                        const T* __restrict__ y,
                                                        const int32_t* __restrict__ num_atoms,
                                                        const double* __restrict__ state,
                                                        uint32_t* status, int A, int nblk) {
    const int b = blockIdx.x / nblk;
    const int i = (blockIdx.x - b * nblk) * kBwdThreads + threadIdx.x;
    if (i >= A) return;
    const int n = clamp_count(num_atoms[b], A, status, false);
    T* g = gx + (size_t)b * 3 * A + 3 * (size_t)i;
    if (i >= n) {
        g[0] = (T)0;
        g[1] = (T)0;
        g[2] = (T)0;
        return;
    }
    const double* s = state + (size_t)b * kState;
    const double scale = (double)grad_rmsd[b] / (s[15] * (double)n);
    const T* xi = x + (size_t)b * 3 * A + 3 * (size_t)i;
    const T* yi = y + (size_t)b * 3 * A + 3 * (size_t)i;
    double dx[3], dy[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        dx[c] = (double)xi[c] - s[c];
        dy[c] = (double)yi[c] - s[3 + c];
    }
    bool bad = false;
#pragma unroll
    for (int c = 0; c < 3; ++c) {  // (U^T dy)_c = sum_k U[k][c] dy_k
        const double uty = s[6 + c] * dy[0] + s[9 + c] * dy[1] + s[12 + c] * dy[2];
        const T v = (T)((dx[c] - uty) * scale);
        bad |= (v != v);
        g[c] = v;
    }
    if (bad) report(status, TPL_RMSD_STATUS_NAN);
}

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

int check(int dtype, int B, int A) {
    if (dtype != TPL_RMSD_F32 && dtype != TPL_RMSD_F64) return TPL_RMSD_ERR_ARG;
    if (B < 0 || A < 0) return TPL_RMSD_ERR_ARG;
    if ((int64_t)3 * A > INT32_MAX || (int64_t)9 * B > INT32_MAX) return TPL_RMSD_ERR_SIZE;
    return TPL_RMSD_OK;
}

}  // namespace

extern "C" {

const char* tpl_rmsd_strerror(int code) {
    switch (code) {
        case TPL_RMSD_OK: return "ok";
        case TPL_RMSD_ERR_ARG: return "invalid argument (null pointer, negative size or unknown dtype)";
        case TPL_RMSD_ERR_LAUNCH: return "kernel launch failed";
        case TPL_RMSD_ERR_SIZE: return "batch * atoms_stride too large";
        default: return "unknown error code";
    }
}

size_t tpl_rmsd_state_bytes(int batch) {
    return batch <= 0 ? 0 : align256((size_t)batch * kState * sizeof(double));
}

int tpl_rmsd_forward(int dtype, const void* input, const void* target, const int32_t* num_atoms,
                     void* rmsd, void* ut, void* state, uint32_t* status, int batch,
                     int atoms_stride, tpl_rmsd_stream_t stream) {
    const int rc = check(dtype, batch, atoms_stride);
    if (rc != TPL_RMSD_OK) return rc;
    if (batch == 0) return TPL_RMSD_OK;
    if (!num_atoms || !rmsd || !ut || (atoms_stride > 0 && (!input || !target)))
        return TPL_RMSD_ERR_ARG;
    if (state && (reinterpret_cast<uintptr_t>(state) & 255)) return TPL_RMSD_ERR_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (dtype == TPL_RMSD_F32)
        hipLaunchKernelGGL(rmsd_fwd<float>, dim3(batch), dim3(kThreads), 0, s,
                           static_cast<const float*>(input), static_cast<const float*>(target),
                           num_atoms, static_cast<float*>(rmsd), static_cast<float*>(ut),
                           static_cast<double*>(state), status, atoms_stride);
    else
        hipLaunchKernelGGL(rmsd_fwd<double>, dim3(batch), dim3(kThreads), 0, s,
                           static_cast<const double*>(input), static_cast<const double*>(target),
                           num_atoms, static_cast<double*>(rmsd), static_cast<double*>(ut),
                           static_cast<double*>(state), status, atoms_stride);
    return hipGetLastError() == hipSuccess ? TPL_RMSD_OK : TPL_RMSD_ERR_LAUNCH;
}

int tpl_rmsd_backward(int dtype, void* grad_input, const void* grad_rmsd, const void* input,
                      const void* target, const int32_t* num_atoms, const void* state,
                      uint32_t* status, int batch, int atoms_stride, tpl_rmsd_stream_t stream) {
    const int rc = check(dtype, batch, atoms_stride);
    if (rc != TPL_RMSD_OK) return rc;
    if (batch == 0 || atoms_stride == 0) return TPL_RMSD_OK;
    if (!grad_input || !grad_rmsd || !input || !target || !num_atoms || !state)
        return TPL_RMSD_ERR_ARG;
    if (reinterpret_cast<uintptr_t>(state) & 255) return TPL_RMSD_ERR_ARG;
    const int nblk = (atoms_stride + kBwdThreads - 1) / kBwdThreads;
    if ((int64_t)nblk * batch > INT32_MAX) return TPL_RMSD_ERR_SIZE;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const dim3 grid(nblk * batch);
    if (dtype == TPL_RMSD_F32)
        hipLaunchKernelGGL(rmsd_bwd<float>, grid, dim3(kBwdThreads), 0, s,
                           static_cast<float*>(grad_input), static_cast<const float*>(grad_rmsd),
                           static_cast<const float*>(input), static_cast<const float*>(target),
                           num_atoms, static_cast<const double*>(state), status, atoms_stride, nblk);
    else
        hipLaunchKernelGGL(rmsd_bwd<double>, grid, dim3(kBwdThreads), 0, s,
                           static_cast<double*>(grad_input), static_cast<const double*>(grad_rmsd),
                           static_cast<const double*>(input), static_cast<const double*>(target),
                           num_atoms, static_cast<const double*>(state), status, atoms_stride, nblk);
    return hipGetLastError() == hipSuccess ? TPL_RMSD_OK : TPL_RMSD_ERR_LAUNCH;
}

}  // extern "C"
