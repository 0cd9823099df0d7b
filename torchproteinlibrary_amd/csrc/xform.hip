// Coordinate transforms for MI355X (gfx950): translate, rotate and centre
// batches of structures, forward and backward.  The C ABI is
// include/tpl_xform.h.
//
// The reference launches <<<atoms_stride, batch_size>>> grids (one block per
// atom index, one thread per structure: TransformCUDAKernels.cu:14-17,
// 34-36,51-53,71-74,100-107), so neighbouring threads touch addresses
// 12*atoms_stride bytes apart, and its centre is a serial loop in one thread
// per structure (:77-86).  Here one thread takes one atom slot of one
// structure (neighbouring threads, neighbouring atoms), every slot past
// num_atoms is written 0, and the centre is a workgroup sum in FP64 per
// structure.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tpl_xform.h"

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ int clamp_count(int n, int A, uint32_t* status, bool report_it) {
    if (n < 0 || n > A) {
        if (report_it && status)
            __hip_atomic_fetch_or(status, TPL_XFORM_STATUS_COUNT, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
        return n < 0 ? 0 : A;
    }
    return n;
}

// out_i = M in_i + t for the atoms of the structure, 0 past num_atoms, with
//   M = I (R null), R, or R^T (transpose); in null means in_i = 0;
//   t = tv[b] * tscale (tv null: 0), tscale = 1 or 1/n (per_n).
// Grid: nblk blocks of kThreads atom slots per structure.
template <typename T>
__global__ void __launch_bounds__(kThreads) xform_affine(T* out, const T* __restrict__ in,
                                                         const T* __restrict__ R, int transpose,
                                                         const T* __restrict__ tv, int per_n,
                                                         const int32_t* __restrict__ num_atoms,
                                                         uint32_t* status, int A, int nblk) {
    const int b = blockIdx.x / nblk;
    const int i = (blockIdx.x - b * nblk) * kThreads + threadIdx.x;
    if (i >= A) return;
    const int n = clamp_count(num_atoms[b], A, status, i == 0);
    T* o = out + (size_t)b * 3 * A + 3 * (size_t)i;
    if (i >= n) {
        o[0] = (T)0;
        o[1] = (T)0;
        o[2] = (T)0;
        return;
    }
    T v[3] = {(T)0, (T)0, (T)0};
    if (in) {
        const T* p = in + (size_t)b * 3 * A + 3 * (size_t)i;
        v[0] = p[0];
        v[1] = p[1];
        v[2] = p[2];
    }
    if (R) {
        const T* m = R + (size_t)b * 9;
        T w[3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
            w[r] = transpose ? m[r] * v[0] + m[3 + r] * v[1] + m[6 + r] * v[2]
                             : m[3 * r] * v[0] + m[3 * r + 1] * v[1] + m[3 * r + 2] * v[2];
#pragma unroll
        for (int r = 0; r < 3; ++r) v[r] = w[r];
    }
    if (tv) {
        const T s = per_n ? (T)(1.0 / (double)n) : (T)1;
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] += tv[3 * (size_t)b + c] * s;
    }
    o[0] = v[0];
    o[1] = v[1];
    o[2] = v[2];
}

// center[b] = (sum of the first n atoms) / n, summed in FP64: each thread over
// a strided share of the atoms, then two levels through LDS.
template <typename T>
__global__ void __launch_bounds__(kThreads) xform_center(T* center, const T* __restrict__ in,
                                                         const int32_t* __restrict__ num_atoms,
                                                         uint32_t* status, int A) {
    constexpr int kParts = 16;
    __shared__ double s_m[3 * kThreads];
    __shared__ double s_p[3 * kParts];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int n = clamp_count(num_atoms[b], A, status, tid == 0);
    const T* p = in + (size_t)b * 3 * A;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int i = tid; i < n; i += kThreads) {
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[c] += (double)p[3 * (size_t)i + c];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) s_m[c * kThreads + tid] = acc[c];
    __syncthreads();
    if (tid < 3 * kParts) {
        const int c = tid / kParts, part = tid - c * kParts;
        double s = 0.0;
        for (int t = part; t < kThreads; t += kParts) s += s_m[c * kThreads + t];
        s_p[c * kParts + part] = s;
    }
    __syncthreads();
    if (tid < 3) {
        double s = 0.0;
#pragma unroll
        for (int part = 0; part < kParts; ++part) s += s_p[tid * kParts + part];
        center[3 * (size_t)b + tid] = (T)(s / (double)n);
    }
}

int check(int dtype, int B, int A) {
    if (dtype != TPL_XFORM_F32 && dtype != TPL_XFORM_F64) return TPL_XFORM_ERR_ARG;
    if (B < 0 || A < 0) return TPL_XFORM_ERR_ARG;
    if ((int64_t)3 * A > INT32_MAX || (int64_t)9 * B > INT32_MAX) return TPL_XFORM_ERR_SIZE;
    return TPL_XFORM_OK;
}

int launch_affine(int dtype, void* out, const void* in, const void* R, int transpose,
                  const void* tv, int per_n, const int32_t* num_atoms, uint32_t* status, int B,
                  int A, tpl_xform_stream_t stream) {
    const int nblk = (A + kThreads - 1) / kThreads;
    if ((int64_t)nblk * B > INT32_MAX) return TPL_XFORM_ERR_SIZE;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (dtype == TPL_XFORM_F32)
        hipLaunchKernelGGL(xform_affine<float>, dim3(nblk * B), dim3(kThreads), 0, s,
                           static_cast<float*>(out), static_cast<const float*>(in),
                           static_cast<const float*>(R), transpose, static_cast<const float*>(tv),
                           per_n, num_atoms, status, A, nblk);
    else
        hipLaunchKernelGGL(xform_affine<double>, dim3(nblk * B), dim3(kThreads), 0, s,
                           static_cast<double*>(out), static_cast<const double*>(in),
                           static_cast<const double*>(R), transpose,
                           static_cast<const double*>(tv), per_n, num_atoms, status, A, nblk);
    return hipGetLastError() == hipSuccess ? TPL_XFORM_OK : TPL_XFORM_ERR_LAUNCH;
}

}  // namespace

extern "C" {

const char* tpl_xform_strerror(int code) {
    switch (code) {
        case TPL_XFORM_OK: return "ok";
        case TPL_XFORM_ERR_ARG: return "invalid argument (null pointer, negative size or unknown dtype)";
        case TPL_XFORM_ERR_LAUNCH: return "kernel launch failed";
        case TPL_XFORM_ERR_SIZE: return "sizes too large";
        default: return "unknown error code";
    }
}

int tpl_xform_translate(int dtype, void* out, const void* in, const void* T,
                        const int32_t* num_atoms, uint32_t* status, int batch, int atoms_stride,
                        tpl_xform_stream_t stream) {
    const int rc = check(dtype, batch, atoms_stride);
    if (rc != TPL_XFORM_OK) return rc;
    if (batch == 0 || atoms_stride == 0) return TPL_XFORM_OK;
    if (!out || !in || !T || !num_atoms) return TPL_XFORM_ERR_ARG;
    return launch_affine(dtype, out, in, nullptr, 0, T, 0, num_atoms, status, batch, atoms_stride,
                         stream);
}

int tpl_xform_translate_backward(int dtype, void* grad_in, const void* grad_out,
                                 const int32_t* num_atoms, uint32_t* status, int batch,
                                 int atoms_stride, tpl_xform_stream_t stream) {
    const int rc = check(dtype, batch, atoms_stride);
    if (rc != TPL_XFORM_OK) return rc;
    if (batch == 0 || atoms_stride == 0) return TPL_XFORM_OK;
    if (!grad_in || !grad_out || !num_atoms) return TPL_XFORM_ERR_ARG;
    return launch_affine(dtype, grad_in, grad_out, nullptr, 0, nullptr, 0, num_atoms, status, batch,
                         atoms_stride, stream);
}

int tpl_xform_rotate(int dtype, void* out, const void* in, const void* R, int transpose,
                     const int32_t* num_atoms, uint32_t* status, int batch, int atoms_stride,
                     tpl_xform_stream_t stream) {
    const int rc = check(dtype, batch, atoms_stride);
    if (rc != TPL_XFORM_OK) return rc;
    if (batch == 0 || atoms_stride == 0) return TPL_XFORM_OK;
    if (!out || !in || !R || !num_atoms) return TPL_XFORM_ERR_ARG;
    return launch_affine(dtype, out, in, R, transpose ? 1 : 0, nullptr, 0, num_atoms, status, batch,
                         atoms_stride, stream);
}

int tpl_xform_center(int dtype, void* center, const void* in, const int32_t* num_atoms,
                     uint32_t* status, int batch, int atoms_stride, tpl_xform_stream_t stream) {
    const int rc = check(dtype, batch, atoms_stride);
    if (rc != TPL_XFORM_OK) return rc;
    if (batch == 0) return TPL_XFORM_OK;
    if (!center || !num_atoms || (atoms_stride > 0 && !in)) return TPL_XFORM_ERR_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (dtype == TPL_XFORM_F32)
        hipLaunchKernelGGL(xform_center<float>, dim3(batch), dim3(kThreads), 0, s,
                           static_cast<float*>(center), static_cast<const float*>(in), num_atoms,
                           status, atoms_stride);
    else
        hipLaunchKernelGGL(xform_center<double>, dim3(batch), dim3(kThreads), 0, s,
                           static_cast<double*>(center), static_cast<const double*>(in), num_atoms,
                           status, atoms_stride);
    return hipGetLastError() == hipSuccess ? TPL_XFORM_OK : TPL_XFORM_ERR_LAUNCH;
}

int tpl_xform_center_backward(int dtype, void* grad_in, const void* grad_center,
                              const int32_t* num_atoms, uint32_t* status, int batch,
                              int atoms_stride, tpl_xform_stream_t stream) {
    const int rc = check(dtype, batch, atoms_stride);
    if (rc != TPL_XFORM_OK) return rc;
    if (batch == 0 || atoms_stride == 0) return TPL_XFORM_OK;
    if (!grad_in || !grad_center || !num_atoms) return TPL_XFORM_ERR_ARG;
    return launch_affine(dtype, grad_in, nullptr, nullptr, 0, grad_center, 1, num_atoms, status,
                         batch, atoms_stride, stream);
}

}  // extern "C"
