// Coordinate transforms for MI355X (gfx950): translate, rotate and centre
// batches of structures, forward and backward.  The C ABI is
// include/tpl_xform.h.
//
// The reference launches <<<atoms_stride, batch_size>>> grids (one block per
// atom index, one thread per structure: TransformCUDAKernels.cu:14-17,
// 34-36,51-53,71-74,100-107), so neighbouring threads touch addresses
// 12*atoms_stride bytes apart, and its centre is a serial loop in one thread
// per structure (:77-86).  Here every global access is a 16-byte one at
// consecutive addresses across a wave: translation and rotation stage a
// wave's chunk of atoms through LDS, the centre's backward writes 16-byte
// blocks directly; every slot past num_atoms is written 0, and the centre is
// a workgroup sum in FP64 per structure.
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

constexpr int kAtoms = 4;  // atoms per 16-byte run in the centre's loads
constexpr int kBlocksPerThread = 3;  // 16-byte blocks per thread in xform_shift

__device__ __forceinline__ bool aligned16(const void* p) {
    return (reinterpret_cast<uintptr_t>(p) & 15u) == 0;
}

// out = in + t (in null: 0) over the structure's 3A values taken flat, 0 past
// num_atoms, with t = tv[b] * tscale (tv null: 0), tscale = 1 or 1/n (per_n).
// The translations and the centre's backward need no atom grouping, so each
// thread moves 16-byte blocks at consecutive addresses across the wave: one
// store instruction covers 1 KB of contiguous output.  Grid: nblk blocks of
// kThreads * kBlocksPerThread 16-byte blocks per structure.
template <typename T>
__global__ void __launch_bounds__(kThreads) xform_shift(T* out, const T* __restrict__ in,
                                                        const T* __restrict__ tv, int per_n,
                                                        const int32_t* __restrict__ num_atoms,
                                                        uint32_t* status, int A, int nblk) {
    constexpr int E = 16 / (int)sizeof(T);  // values per 16-byte block
    const int b = blockIdx.x / nblk;
    const int64_t blk0 = (int64_t)(blockIdx.x - b * nblk) * kThreads * kBlocksPerThread;
    const int n = clamp_count(num_atoms[b], A, status, blk0 == 0 && threadIdx.x == 0);
    const int64_t vals = 3 * (int64_t)A, valid = 3 * (int64_t)n;
    T* o = out + (size_t)b * vals;
    const T* p = in ? in + (size_t)b * vals : nullptr;
    T t[3] = {(T)0, (T)0, (T)0};
    if (tv && n > 0) {
        const T sc = per_n ? (T)(1.0 / (double)n) : (T)1;
#pragma unroll
        for (int c = 0; c < 3; ++c) t[c] = tv[3 * (size_t)b + c] * sc;
    }
    const bool vec = aligned16(o) && (!p || aligned16(p));
#pragma unroll
    for (int r = 0; r < kBlocksPerThread; ++r) {
        const int64_t v0 = ((int64_t)blk0 + r * kThreads + threadIdx.x) * E;
        if (v0 >= vals) break;
        T v[E];
        if (vec && v0 + E <= vals) {
            if (p && v0 < valid) {
                const uint4 u = *reinterpret_cast<const uint4*>(p + v0);
                __builtin_memcpy(v, &u, 16);
            } else {
#pragma unroll
                for (int e = 0; e < E; ++e) v[e] = (T)0;
            }
            // v0 < 3A <= INT32_MAX here: a 32-bit remainder is exact (a 64-bit
            // one is a long software sequence on the GPU)
            int c = (int)((uint32_t)v0 % 3u);
#pragma unroll
            for (int e = 0; e < E; ++e) {
                v[e] = v0 + e < valid ? v[e] + t[c] : (T)0;
                c = c == 2 ? 0 : c + 1;
            }
            uint4 u;
            __builtin_memcpy(&u, v, 16);
            *reinterpret_cast<uint4*>(o + v0) = u;
        } else {
            for (int e = 0; e < E && v0 + e < vals; ++e) {
                const int64_t i = v0 + e;
                o[i] = i < valid ? (p ? p[i] : (T)0) + t[(uint32_t)i % 3u] : (T)0;
            }
        }
    }
}

// out_i = M in_i + t for the atoms of the structure (M = R, R^T with
// transpose, or I when R is null; t = tv[b], or 0 when tv is null), 0 past
// num_atoms, with every global access a 16-byte one at consecutive addresses
// across the wave.  Each wave takes a chunk of kWaveAtoms atoms
// (3 * kWaveAtoms values): it loads the chunk into its own slice of LDS with
// coalesced 16-byte loads, each lane rotates four whole atoms there, and the
// wave stores the chunk back with coalesced 16-byte stores.  The slice is the
// wave's alone, so the wave needs no workgroup barrier; the wavefront fences
// and the wave barrier keep the LDS writes before the reads.  A chunk that
// runs past the structure's stride or starts unaligned goes atom by atom.
// Grid: nblk blocks of 4 chunks per structure.
constexpr int kWaveAtoms = 256;

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
__device__ __forceinline__ void rotate_atom(const T m[9], int transpose, const T v[3], T w[3]) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
        w[r] = transpose ? m[r] * v[0] + m[3 + r] * v[1] + m[6 + r] * v[2]
                         : m[3 * r] * v[0] + m[3 * r + 1] * v[1] + m[3 * r + 2] * v[2];
}

template <typename T>
__global__ void __launch_bounds__(kThreads) xform_waves(T* out, const T* __restrict__ in,
                                                        const T* __restrict__ R, int transpose,
                                                        const T* __restrict__ tv,
                                                        const int32_t* __restrict__ num_atoms,
                                                        uint32_t* status, int A, int nblk) {
    constexpr int E = 16 / (int)sizeof(T);  // values per 16-byte access
    constexpr int V = 3 * kWaveAtoms;       // values per chunk
    constexpr int kPerLane = kWaveAtoms / 64;
    __shared__ T s_buf[kThreads / 64][V];
    const int b = blockIdx.x / nblk;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int a0 = ((blockIdx.x - b * nblk) * (kThreads / 64) + wave) * kWaveAtoms;
    if (a0 >= A) return;  // the whole wave: no workgroup barrier follows
    const int n = clamp_count(num_atoms[b], A, status, a0 == 0 && lane == 0);
    T* o = out + (size_t)b * 3 * A;
    const T* p = in + (size_t)b * 3 * A;
    T m[9], t[3] = {(T)0, (T)0, (T)0};
#pragma unroll
    for (int k = 0; k < 9; ++k) m[k] = R ? R[(size_t)b * 9 + k] : (T)(k % 4 == 0 ? 1 : 0);
    if (tv) {
#pragma unroll
        for (int c = 0; c < 3; ++c) t[c] = tv[3 * (size_t)b + c];
    }
    if (a0 + kWaveAtoms <= A && aligned16(o + 3 * (size_t)a0) && aligned16(p + 3 * (size_t)a0)) {
        T* buf = s_buf[wave];
        if (a0 < n) {
#pragma unroll
            for (int j = 0; j < V / (64 * E); ++j) {
                const int idx = (j * 64 + lane) * E;
                const uint4 u = *reinterpret_cast<const uint4*>(p + 3 * (size_t)a0 + idx);
                *reinterpret_cast<uint4*>(buf + idx) = u;
            }
        }
        wave_lds_sync();
#pragma unroll
        for (int k = 0; k < kPerLane; ++k) {
            const int q = kPerLane * lane + k;
            T v[3] = {buf[3 * q], buf[3 * q + 1], buf[3 * q + 2]}, w[3] = {(T)0, (T)0, (T)0};
            if (a0 + q < n) {
                if (R) rotate_atom(m, transpose, v, w);
                else w[0] = v[0], w[1] = v[1], w[2] = v[2];
#pragma unroll
                for (int c = 0; c < 3; ++c) w[c] += t[c];
            }
            buf[3 * q] = w[0];
            buf[3 * q + 1] = w[1];
            buf[3 * q + 2] = w[2];
        }
        wave_lds_sync();
#pragma unroll
        for (int j = 0; j < V / (64 * E); ++j) {
            const int idx = (j * 64 + lane) * E;
            *reinterpret_cast<uint4*>(o + 3 * (size_t)a0 + idx) = *reinterpret_cast<const uint4*>(buf + idx);
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < kPerLane; ++k) {
        const int i = a0 + kPerLane * lane + k;
        if (i >= A) break;
        T v[3] = {(T)0, (T)0, (T)0}, w[3] = {(T)0, (T)0, (T)0};
        if (i < n) {
            v[0] = p[3 * (size_t)i];
            v[1] = p[3 * (size_t)i + 1];
            v[2] = p[3 * (size_t)i + 2];
            if (R) rotate_atom(m, transpose, v, w);
            else w[0] = v[0], w[1] = v[1], w[2] = v[2];
#pragma unroll
            for (int c = 0; c < 3; ++c) w[c] += t[c];
        }
        o[3 * (size_t)i] = w[0];
        o[3 * (size_t)i + 1] = w[1];
        o[3 * (size_t)i + 2] = w[2];
    }
}

// center[b] = (sum of the first n atoms) / n, summed in FP64: each thread over
// a strided share of runs of kAtoms atoms (16-byte loads when the structure
// starts 16-byte aligned), then two levels through LDS.
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
    constexpr int kPer16 = 16 / (int)sizeof(T);
    double acc[3] = {0.0, 0.0, 0.0};
    const bool vec = aligned16(p);
    for (int a0 = tid * kAtoms; a0 < n; a0 += kThreads * kAtoms) {
        if (vec && a0 + kAtoms <= n) {
            T v[3 * kAtoms];
#pragma unroll
            for (int k = 0; k < 3 * kAtoms; k += kPer16) {
                const uint4 u = *reinterpret_cast<const uint4*>(p + 3 * (size_t)a0 + k);
                __builtin_memcpy(v + k, &u, 16);
            }
#pragma unroll
            for (int k = 0; k < 3 * kAtoms; ++k) acc[k % 3] += (double)v[k];
        } else {
            for (int i = a0; i < n && i < a0 + kAtoms; ++i) {
#pragma unroll
                for (int c = 0; c < 3; ++c) acc[c] += (double)p[3 * (size_t)i + c];
            }
        }
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
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (in && !per_n) {
        const int64_t per = (int64_t)(kThreads / 64) * kWaveAtoms;
        const int64_t nb = ((int64_t)A + per - 1) / per;
        if (nb * B > INT32_MAX) return TPL_XFORM_ERR_SIZE;
        const int nblk = (int)nb;
        if (dtype == TPL_XFORM_F32)
            hipLaunchKernelGGL(xform_waves<float>, dim3(nblk * B), dim3(kThreads), 0, s,
                               static_cast<float*>(out), static_cast<const float*>(in),
                               static_cast<const float*>(R), transpose, static_cast<const float*>(tv),
                               num_atoms, status, A, nblk);
        else
            hipLaunchKernelGGL(xform_waves<double>, dim3(nblk * B), dim3(kThreads), 0, s,
                               static_cast<double*>(out), static_cast<const double*>(in),
                               static_cast<const double*>(R), transpose, static_cast<const double*>(tv),
                               num_atoms, status, A, nblk);
        return hipGetLastError() == hipSuccess ? TPL_XFORM_OK : TPL_XFORM_ERR_LAUNCH;
    }
    if (!R) {
        const int64_t per = (int64_t)kThreads * kBlocksPerThread * (dtype == TPL_XFORM_F32 ? 4 : 2);
        const int64_t nb = (3 * (int64_t)A + per - 1) / per;
        if (nb * B > INT32_MAX) return TPL_XFORM_ERR_SIZE;
        const int nblk = (int)nb;
        if (dtype == TPL_XFORM_F32)
            hipLaunchKernelGGL(xform_shift<float>, dim3(nblk * B), dim3(kThreads), 0, s,
                               static_cast<float*>(out), static_cast<const float*>(in),
                               static_cast<const float*>(tv), per_n, num_atoms, status, A, nblk);
        else
            hipLaunchKernelGGL(xform_shift<double>, dim3(nblk * B), dim3(kThreads), 0, s,
                               static_cast<double*>(out), static_cast<const double*>(in),
                               static_cast<const double*>(tv), per_n, num_atoms, status, A, nblk);
        return hipGetLastError() == hipSuccess ? TPL_XFORM_OK : TPL_XFORM_ERR_LAUNCH;
    }
    return TPL_XFORM_ERR_ARG;  // a rotation needs an input
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
