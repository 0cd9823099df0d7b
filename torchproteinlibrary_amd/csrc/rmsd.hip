// Coords2RMSD for MI355X (gfx950): minimal RMSD over rigid rotations and its
// gradient, each direction one kernel launch.  The C ABI is include/tpl_rmsd.h.
//
// The layer follows the reference's method (quaternion eigenproblem of the
// centred correlation matrix; coords2rmsd_interface.cpp:12-49,52-111,
// RMSDKernels.cu:5-39) with everything on the device:
//   forward, one workgroup per structure:
//     1. FP64 moments of the atoms relative to the structure's first atom
//        (sum x, sum y, sum x y^T, sum |x|^2, sum |y|^2), four atoms per
//        thread and pass, summed over the workgroup in two levels through LDS;
//     2. lane 0 centres them, builds the symmetric 4x4 matrix F of the
//        centred correlation R (the reference's element layout,
//        coords2rmsd_interface.cpp:153-158), takes its largest eigenpair
//        (rmsd_math.h: Newton on the characteristic quartic, eigenvector from
//        the adjugate; a shifted adjugate for a double root), forms U from the quaternion
//        (same formula, :35-45) and rmsd = sqrt((R2 - 2 |lambda| + 1e-10) / n)
//        (:182);
//   backward, one thread per atom: (x - cx - U^T (y - cy)) * g / (rmsd * n)
//     (Coords2RMSD.py:47-56), zero past num_atoms.
// Shifting by the first atom keeps the raw moments at the size of the
// structure, so centring them afterwards loses no digits to a large offset
// from the origin.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "rmsd_math.h"
#include "tpl_rmsd.h"

namespace {

constexpr double kEpsLimit = 1e-10;  // coords2rmsd_interface.cpp:10
constexpr int kThreads = 256;        // forward workgroup: 4 waves
constexpr int kBwdThreads = 256;     // backward: one thread per atom
constexpr int kState = 16;           // doubles per structure: cx[3] cy[3] U[9] rmsd
constexpr int kMoments = 17;         // sx[3] sy[3] sxy[9] sxx syy

// Atoms a0 .. a0+3 of one structure (xyz interleaved) as 12 values: three
// 16-byte loads per 4 floats (six for doubles) when the run is aligned and
// whole, else element loads; atoms at or past `lim` read as 0.
template <typename T>
__device__ __forceinline__ void load4(const T* p, int a0, int lim, bool vec, double v[12]) {
    if (vec && a0 + 4 <= lim) {
        constexpr int E = 16 / (int)sizeof(T);  // elements per 16-byte load
        T t[12];
#pragma unroll
        for (int k = 0; k < 12; k += E) {
            const uint4 u = *reinterpret_cast<const uint4*>(p + 3 * (size_t)a0 + k);
            __builtin_memcpy(t + k, &u, 16);
        }
#pragma unroll
        for (int k = 0; k < 12; ++k) v[k] = (double)t[k];
    } else {
#pragma unroll
        for (int k = 0; k < 12; ++k) v[k] = (a0 + k / 3 < lim) ? (double)p[3 * (size_t)a0 + k] : 0.0;
    }
}

__device__ __forceinline__ bool aligned16(const void* p) {
    return (reinterpret_cast<uintptr_t>(p) & 15) == 0;
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

template <typename T>
__global__ void __launch_bounds__(kThreads) rmsd_fwd(const T* __restrict__ x, const T* __restrict__ y,
                                                     const int32_t* __restrict__ num_atoms,
                                                     T* rmsd_out, T* ut_out, double* state,
                                                     uint32_t* status, int A) {
    // Moment sums over the workgroup: every thread's 17 partial sums (laid
    // out [moment][thread]), then kParts partial sums per moment, then the
    // totals.  A cross-lane butterfly over 17 doubles measured 5 us per
    // workgroup, bound by the LDS crossbar at 16 waves per CU; this takes
    // three barriers and plain LDS reads.
    constexpr int kParts = kThreads / kMoments;  // 15 threads per moment in the middle level
    __shared__ double s_m[kMoments * kThreads];
    __shared__ double s_p[kMoments * kParts];
    __shared__ double s_tot[kMoments];
    const int b = blockIdx.x, tid = threadIdx.x;
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
    // Four consecutive atoms per thread and pass: one 48-byte run of each
    // structure per lane, loaded in 16-byte pieces.
    const bool vec = aligned16(xb) && aligned16(yb);
    for (int a0 = 4 * tid; a0 < n; a0 += 4 * kThreads) {
        double xv[12], yv[12];
        load4(xb, a0, n, vec, xv);
        load4(yb, a0, n, vec, yv);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const bool in = a0 + j < n;
            double dx[3], dy[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                dx[c] = in ? xv[3 * j + c] - x0[c] : 0.0;
                dy[c] = in ? yv[3 * j + c] - y0[c] : 0.0;
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
    }
#pragma unroll
    for (int e = 0; e < kMoments; ++e) s_m[e * kThreads + tid] = m[e];
    __syncthreads();
    if (tid < kMoments * kParts) {
        const int e = tid / kParts, part = tid - e * kParts;
        double acc = 0.0;
        for (int t = part; t < kThreads; t += kParts) acc += s_m[e * kThreads + t];
        s_p[e * kParts + part] = acc;
    }
    __syncthreads();
    if (tid < kMoments) {
        double acc = 0.0;
#pragma unroll
        for (int part = 0; part < kParts; ++part) acc += s_p[tid * kParts + part];
        s_tot[tid] = acc;
    }
    __syncthreads();
    if (tid != 0) return;
#pragma unroll
    for (int e = 0; e < kMoments; ++e) m[e] = s_tot[e];
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
    double F[16], U[9];
    rmsd_math::f_matrix(R, F);
    const rmsd_math::Eigenpair eig = rmsd_math::max_eigenpair(F, 0.5 * r2);
    const double lam = eig.lam;
    rmsd_math::u_from_q(eig, U);
    // r2 - 2|lambda| is >= 0 in exact arithmetic; for (nearly) identical
    // structures its FP64 rounding (~1e-16 r2) can make it negative, which
    // the reference's formula turns into NaN.  Clamped at 0 here; a NaN (from
    // NaN coordinates) stays NaN and is reported.
    const double dev = r2 - 2.0 * fabs(lam);
    const double rmsd = sqrt(((dev < 0.0 ? 0.0 : dev) + kEpsLimit) / nd);
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
