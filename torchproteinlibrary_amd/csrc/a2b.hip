// a2b.hip -- Angles2Backbone forward/backward for MI355X (gfx950).
//
// Reference path: TorchProteinLibrary ReducedModel.Angles2Backbone
// (Layers/ReducedModel/cBackboneProteinCUDAKernels.cu, cBackboneProteinCPUKernels.cpp).
// Backbone atom i >= 1 of a chain sits at the origin of frame
//   A_i = A_{i-1} * B(theta_i, beta_i, d_i),  A_0 = I,
//   B(theta, beta, d) = Ry(beta) Tx(d) Rx(theta)
// (cBackboneProteinCPUKernels.cpp:36-58, Math/cMatrix44.cpp:47-52).  Residue r
// owns the three transforms (phi_r, psi_r, omega_r) that place CA_r, C_r and
// N_{r+1}, so every lane reads its own angles and writes its own gradients.
//
// Chunk kernels (a2b_fwd_chunk / a2b_bwd_chunk), FP64 arithmetic:
//   * a chunk is 64*R consecutive residues of one chain and one wave takes
//     one chunk; lane l owns R of them.  The slots of all chains form one
//     flat list, chunk-major inside XCD stripes; slots wholly past a chain
//     write its padding zeros, the others take the chain's chunks through a
//     per-chain ticket, so work spreads over every SIMD whatever the mix of
//     lengths;
//   * forward: lane l composes its R residues serially (the structured 3x4
//     step below), the 64 lane aggregates are scanned with DPP (row_shr
//     1/2/4/8, row_bcast 15/31), and chunks of one chain are chained by a
//     wave-parallel decoupled look-back (each lane polls one earlier chunk;
//     the ordered product of their frames is another wave scan).  The frame
//     at every backward-chunk boundary is saved as a checkpoint;
//   * backward: the O(L) adjoint
//       dL/dtheta_j = u_j . (T_j - x_j x F_j),
//       F_j = sum_{i>=j} g_i,  T_j = sum_{i>=j} x_i x g_i,
//     with u_j the unit bond vector ending at atom j.  Each chunk recomputes
//     its frames from its angles and its checkpoint (no coordinates or
//     per-atom state are read back), sums (x x g, g) over its atoms, takes
//     the suffix sums over lanes (DPP) and later chunks (look-back), and
//     walks its atoms once more for the gradients.
// Chain kernels (a2b_fwd_chain / a2b_bwd_chain): the same chunk steps with
// one workgroup per chain of at most 4 chunks, the chunks meeting in LDS
// instead of through the look-back; used when every chain's workgroup is
// resident at once.
// Plain kernels (a2b_fwd_plain / a2b_bwd_plain) restate the reference GPU
// algorithm without the O(L^2) derivative tensor: they are the baseline.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "sincos_tab.h"
#include "tpl_a2b.h"

namespace {

// ----------------------------------------------------------------- geometry
// Bond lengths and beta = pi - bond angle (cBackboneProteinCPUKernels.cpp:6-12).
// cos/sin(beta) are the IEEE double values of cos(pi - a), sin(pi - a).
// Index 0: N->CA (phi), 1: CA->C (psi), 2: C->N (omega).
struct Bond {
    double cb, sb, d;
};
__device__ __host__ constexpr Bond bond(int k) {
    return k == 0 ? Bond{0.3600333805869535, 0.932939421861425, 1.460}
         : k == 1 ? Bond{0.470805585969763, 0.8822369864269283, 1.525}
                  : Bond{0.5208135473026831, 0.8536704568778257, 1.330};
}

// A frame is a 3x4 affine map stored by columns: f[0..2] = R e_x, f[3..5] =
// R e_y, f[6..8] = R e_z, f[9..11] = t.
constexpr int kF = 12;

__device__ __forceinline__ double ident(int e) {
    return (e == 0 || e == 4 || e == 8) ? 1.0 : 0.0;
}

__device__ __forceinline__ void frame_identity(double f[kF]) {
#pragma unroll
    for (int e = 0; e < kF; ++e) f[e] = ident(e);
}

// f <- f * B(theta, beta_K, d_K) given (cos theta, sin theta).
//   f.R Ry(beta): c0' = cb c0 - sb c2, c1' = c1, c2' = sb c0 + cb c2
//   * Rx(theta):  c1'' = c c1' + s c2', c2'' = c c2' - s c1'
//   t' = t + d c0'
template <int K>
__device__ __forceinline__ void frame_step(double f[kF], double c, double s) {
    constexpr Bond bd = bond(K);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double a0 = f[i], a1 = f[3 + i], a2 = f[6 + i];
        const double n0 = bd.cb * a0 - bd.sb * a2;
        const double p2 = bd.sb * a0 + bd.cb * a2;
        f[i] = n0;
        f[3 + i] = c * a1 + s * p2;
        f[6 + i] = c * p2 - s * a1;
        f[9 + i] = fma(bd.d, n0, f[9 + i]);
    }
}

template <typename P>
__device__ __forceinline__ void frame_lmul(const P& M, double f[kF]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const double a = f[3 * j], b = f[3 * j + 1], c = f[3 * j + 2];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double v = M[i] * a + M[3 + i] * b + M[6 + i] * c;
            f[3 * j + i] = (j == 3) ? v + M[9 + i] : v;
        }
    }
}

// out <- f * M, written column by column to memory (f in registers, M and
// out in LDS or global memory); needs no full temporary frame.
__device__ __forceinline__ void frame_rmul_to(const double f[kF], const double* M, double* out) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const double a = M[3 * j], b = M[3 * j + 1], c = M[3 * j + 2];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double v = f[i] * a + f[3 + i] * b + f[6 + i] * c;
            out[3 * j + i] = (j == 3) ? v + f[9 + i] : v;
        }
    }
}


__device__ __forceinline__ double dbl_from(int lo, int hi) {
    return __hiloint2double(hi, lo);
}

// DPP move of a double with `old` for lanes that receive nothing.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double v, double old) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(v), CTRL,
                                               ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(v), CTRL,
                                               ROW_MASK, 0xf, false);
    return dbl_from(lo, hi);
}

__device__ __forceinline__ double read_lane(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return dbl_from(lo, hi);
}

// DPP controls (GFX9 encoding).
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143, kWaveShr1 = 0x138;

// DPP move of a double where lanes that receive nothing read 0
// (bound_ctrl); needs no `old` value, hence no register initialisation.
template <int CTRL>
__device__ __forceinline__ double dpp_move_zero(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xf, 0xf, true);
    return dbl_from(lo, hi);
}

// One Hillis-Steele step: f <- (frame of the source lane) * f; lanes that
// receive nothing compose with the identity.  For the row shifts (every row
// enabled) the zero elements of the identity come from bound_ctrl and only
// the high words of the three diagonal 1.0s need an `old` value.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void frame_scan_step(double f[kF]) {
    double g[kF];
    if constexpr (ROW_MASK == 0xf) {
#pragma unroll
        for (int e = 0; e < kF; ++e) {
            if (ident(e) == 0.0) {
                g[e] = dpp_move_zero<CTRL>(f[e]);
            } else {
                const int lo = __builtin_amdgcn_mov_dpp(__double2loint(f[e]), CTRL, 0xf, 0xf, true);
                const int hi = __builtin_amdgcn_update_dpp(__double2hiint(1.0), __double2hiint(f[e]),
                                                           CTRL, 0xf, 0xf, false);
                g[e] = dbl_from(lo, hi);
            }
        }
    } else {
#pragma unroll
        for (int e = 0; e < kF; ++e) g[e] = dpp_move<CTRL, ROW_MASK>(f[e], ident(e));
    }
    frame_lmul(g, f);
}

// Inclusive scan of frames over the 64 lanes of a wave, in lane order
// (lane l ends with f_0 * f_1 * ... * f_l).  All 64 lanes must be active.
__device__ __forceinline__ void wave_scan_frames(double f[kF]) {
    frame_scan_step<kRowShr1, 0xf>(f);
    frame_scan_step<kRowShr2, 0xf>(f);
    frame_scan_step<kRowShr4, 0xf>(f);
    frame_scan_step<kRowShr8, 0xf>(f);
    frame_scan_step<kRowBcast15, 0xa>(f);
    frame_scan_step<kRowBcast31, 0xc>(f);
}

// Exclusive version from an inclusive scan, in place: lane l gets lane l-1
// (lane 0: I).
__device__ __forceinline__ void wave_shift_frames(double f[kF]) {
#pragma unroll
    for (int e = 0; e < kF; ++e) f[e] = dpp_move<kWaveShr1, 0xf>(f[e], ident(e));
}

// Inclusive prefix sum over the wave of N doubles (row shifts zero-fill
// through bound_ctrl).
template <int N>
__device__ __forceinline__ void wave_scan_sum(double v[N]) {
#pragma unroll
    for (int e = 0; e < N; ++e) {
        double x = v[e];
        x += dpp_move_zero<kRowShr1>(x);
        x += dpp_move_zero<kRowShr2>(x);
        x += dpp_move_zero<kRowShr4>(x);
        x += dpp_move_zero<kRowShr8>(x);
        x += dpp_move<kRowBcast15, 0xa>(x, 0.0);
        x += dpp_move<kRowBcast31, 0xc>(x, 0.0);
        v[e] = x;
    }
}

// --------------------------------------------------------------- I/O helpers
__device__ __forceinline__ void set_status(uint32_t* status, uint32_t bits) {
    if (status) __hip_atomic_fetch_or(status, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Cross-workgroup look-back traffic.  On gfx950 an agent-scope release or
// acquire fence writes back or invalidates the whole L2 of the XCD (the L2s of
// the eight XCDs are not coherent with each other), which the streaming
// output of every other workgroup makes expensive.  So the look-back never
// fences at agent scope: tickets, done counters and record words move with
// relaxed agent-scope atomics, which are coherent at device scope by
// themselves, and a record needs no order between its words (each word
// shows by itself whether it is written).  The one ordering point, a wave's
// record stores before its done count, waits with s_waitcnt vmcnt(0)
// (vector memory operations complete), plus a workgroup-scope fence that
// stops the compiler from reordering.
constexpr int kWaitVmcnt0 = 0 | (0x7 << 4) | (0xf << 8);  // vmcnt(0); expcnt, lgkmcnt: no wait

__device__ __forceinline__ void word_store(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void record_store(uint64_t* p, uint64_t v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t record_load(const uint64_t* p) {
    return __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_inc(uint32_t* p) {
    return __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------------------ lane I/O
// A lane moves a run of N consecutive elements of T between global memory
// and registers.  Elements at or past `cnt` read as 0 and are not written.  A
// run that is whole (cnt >= N) and U-byte aligned moves in U-byte accesses;
// U = the widest of 16/8/4 bytes dividing the run, fixed per kernel, and every
// lane of one row shares the alignment, so the branch is uniform there.
// (Runs here are whole multiples of 2 bytes.)
template <int U>
struct Unit;
template <>
struct Unit<2> {
    using type = uint16_t;
};
template <>
struct Unit<4> {
    using type = uint32_t;
};
template <>
struct Unit<8> {
    using type = uint2;
};
template <>
struct Unit<16> {
    using type = uint4;
};
template <typename T, int N>
constexpr int unit_bytes() {
    constexpr int b = N * (int)sizeof(T);
    return b % 16 == 0 ? 16 : (b % 8 == 0 ? 8 : (b % 4 == 0 ? 4 : 2));
}

// avail: elements at p that may be read (the run's room in its row); the
// whole run is read when avail >= N, and elements at or past cnt are then
// zeroed with selects, so a lane that ends its chain takes no branch.
template <typename T, int N>
__device__ __forceinline__ void lane_load(const T* p, int cnt, T (&v)[N], int avail = -1) {
    constexpr int U = unit_bytes<T, N>();
    using V = typename Unit<U>::type;
    if (avail < 0) avail = cnt;
    if (avail >= N && (reinterpret_cast<uintptr_t>(p) % U) == 0) {
        V w[N * sizeof(T) / U];
#pragma unroll
        for (int i = 0; i < (int)(N * sizeof(T) / U); ++i) w[i] = reinterpret_cast<const V*>(p)[i];
        __builtin_memcpy(v, w, sizeof(v));
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = i < cnt ? v[i] : (T)0;
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = i < cnt ? p[i] : (T)0;
    }
}

// Two runs of the same shape (e.g. coordinates and their gradients) under
// one alignment test, so their loads are in flight together.
template <typename T, int N>
__device__ __forceinline__ void lane_load2(const T* p, const T* q, int cnt, int avail, T (&v)[N],
                                           T (&w)[N]) {
    constexpr int U = unit_bytes<T, N>();
    using V = typename Unit<U>::type;
    const bool vec = avail >= N && ((reinterpret_cast<uintptr_t>(p) | reinterpret_cast<uintptr_t>(q)) % U) == 0;
    if (vec) {
        V a[N * sizeof(T) / U], c[N * sizeof(T) / U];
#pragma unroll
        for (int i = 0; i < (int)(N * sizeof(T) / U); ++i) {
            a[i] = reinterpret_cast<const V*>(p)[i];
            c[i] = reinterpret_cast<const V*>(q)[i];
        }
        __builtin_memcpy(v, a, sizeof(v));
        __builtin_memcpy(w, c, sizeof(w));
#pragma unroll
        for (int i = 0; i < N; ++i) {
            v[i] = i < cnt ? v[i] : (T)0;
            w[i] = i < cnt ? w[i] : (T)0;
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            v[i] = i < cnt ? p[i] : (T)0;
            w[i] = i < cnt ? q[i] : (T)0;
        }
    }
}

// The lane's R angles of the three rows (stride L elements) under one
// alignment test, so the three loads are in flight together.
template <typename T, int R>
__device__ __forceinline__ void lane_load_rows(const T* p, int L, int cnt, int avail, T (&v)[3][R]) {
    constexpr int U = unit_bytes<T, R>();
    using V = typename Unit<U>::type;
    const bool vec = avail >= R && (reinterpret_cast<uintptr_t>(p) % U) == 0 &&
                     ((size_t)L * sizeof(T)) % U == 0;
    if (vec) {
        V a[3][R * sizeof(T) / U];
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int i = 0; i < (int)(R * sizeof(T) / U); ++i) a[k][i] = reinterpret_cast<const V*>(p + (size_t)k * L)[i];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            __builtin_memcpy(v[k], a[k], sizeof(v[k]));
#pragma unroll
            for (int i = 0; i < R; ++i) v[k][i] = i < cnt ? v[k][i] : (T)0;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int i = 0; i < R; ++i) v[k][i] = i < cnt ? p[(size_t)k * L + i] : (T)0;
    }
}

// Vector types of 4, 8 and 16 bytes that __builtin_nontemporal_store takes.
template <int U>
struct StreamUnit;
template <>
struct StreamUnit<2> {
    using type = uint16_t;
};
template <>
struct StreamUnit<4> {
    using type = uint32_t;
};
template <>
struct StreamUnit<8> {
    typedef unsigned int type __attribute__((ext_vector_type(2)));
};
template <>
struct StreamUnit<16> {
    typedef unsigned int type __attribute__((ext_vector_type(4)));
};

// Whole aligned runs go out as nontemporal (streaming) stores: the outputs
// are never read back by the kernel that writes them.
template <typename T, int N>
__device__ __forceinline__ void lane_store(T* p, int cnt, const T (&v)[N]) {
    constexpr int U = unit_bytes<T, N>();
    using V = typename StreamUnit<U>::type;
    if (cnt >= N && (reinterpret_cast<uintptr_t>(p) % U) == 0) {
        V w[N * sizeof(T) / U];
        __builtin_memcpy(w, v, sizeof(v));
#pragma unroll
        for (int i = 0; i < (int)(N * sizeof(T) / U); ++i)
            __builtin_nontemporal_store(w[i], reinterpret_cast<V*>(p) + i);
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (i < cnt) p[i] = v[i];
    }
}

template <typename T, int N>
__device__ __forceinline__ void lane_store_zero(T* p, int cnt) {
    T z[N];
#pragma unroll
    for (int i = 0; i < N; ++i) z[i] = (T)0;
    lane_store<T, N>(p, cnt, z);
}

__device__ __forceinline__ void affine(const double E[kF], double p0, double p1, double p2,
                                       double x[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) x[i] = E[i] * p0 + E[3 + i] * p1 + E[6 + i] * p2 + E[9 + i];
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Wave-wide store of the run dst[0, count) in global memory: element f <
// nsrc is s_w[f], the wave's LDS staging in output order; the rest is 0.  The
// aligned body moves in 16-byte LDS reads and 16-byte global stores,
// lane-consecutive, so every store instruction of the wave writes 1 KiB.  A
// wave with nothing staged (nsrc == 0, uniform) reads nothing.  The body's
// stores are nontemporal (streaming): this kernel never reads the
// coordinates back, and the forward at B=1024, L=1000 measured 20.0 us
// with them against 22.0 us with ordinary stores.
template <typename T>
__device__ __forceinline__ void wave_store_run(T* dst, int count, int nsrc, const T* s_w, int lane) {
    constexpr int V = 16 / (int)sizeof(T);
    const bool staged = nsrc > 0;
    const int mis = (int)((reinterpret_cast<uintptr_t>(dst) / sizeof(T)) % V);
    const int head = min((V - mis) % V, count);
    const int nvec = (count - head) / V;
    const int tail0 = head + nvec * V;
    if (lane < head) dst[lane] = (staged && lane < nsrc) ? s_w[lane] : (T)0;
    for (int v = lane; v < nvec; v += 64) {
        const int f0 = head + v * V;
        T w[V];
        if (staged) {
            if (head == 0) {  // uniform: 16-byte aligned in LDS too
                const uint4 u = *reinterpret_cast<const uint4*>(s_w + f0);
                __builtin_memcpy(w, &u, sizeof(u));
            } else {
#pragma unroll
                for (int e = 0; e < V; ++e) w[e] = s_w[f0 + e];
            }
#pragma unroll
            for (int e = 0; e < V; ++e) w[e] = f0 + e < nsrc ? w[e] : (T)0;
        } else {
#pragma unroll
            for (int e = 0; e < V; ++e) w[e] = (T)0;
        }
        u32x4 u;
        __builtin_memcpy(&u, w, sizeof(u));
        __builtin_nontemporal_store(u, reinterpret_cast<u32x4*>(dst + f0));
    }
    if (tail0 + lane < count) dst[tail0 + lane] = (staged && tail0 + lane < nsrc) ? s_w[tail0 + lane] : (T)0;
}

// Wave-scope LDS ordering: the wave's earlier LDS writes are visible to its
// later LDS reads (the LDS serves one wave's instructions in order; this
// keeps the compiler from reordering across the point).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}


// ------------------------------------------------------------ chunk kernels
// Work decomposition.  A chunk is C = 64*R consecutive residues of one chain
// (lane l owns residues r0 + l*R .. r0 + l*R + R-1, i.e. chain atoms
// 3(r0 + lR) .. 3(r0 + lR) + 3R - 1).  One workgroup of one wave takes one
// slot (b, s), s < nc = ceil(L / C).  Workgroup w runs on XCD w mod 8 (the
// dispatcher's round-robin) and takes a chain b = x, x + 8, x + 16, ... of
// stripe x = w mod 8, so all slots of a chain share one XCD; slot_at below
// says which slots.
// With n residues a chain has nreal = ceil(n / C) chunks.  Slots s >= nreal
// only write the zeros of their span of the outputs.  The nreal slots below
// take the chain's chunks from a per-chain ticket counter, ascending in the
// forward and descending in the backward, so a chunk waits only on chunks
// whose waves are already running, whatever order workgroups start in.  A
// chain of one chunk takes no ticket and publishes nothing.
//
// Look-back records (a chunk's aggregate, or a group boundary's inclusive
// value; one record per chunk) live in the caller's sync buffer and carry
// their own "written" mark: a record word is the FP64 bit pattern of its
// value with bit 0 set, so no written word is zero, and a reader that sees
// every word of a record nonzero has the whole record (each word is written
// once per call and moves as one 64-bit access).  Setting bit 0 changes a
// value by at most one unit in the last place, the same way on every run.
// So a publisher issues its stores and goes on, and a reader polls the
// record itself: one memory round trip between them, and no flag.
// The buffer cleans up after itself: each wave that took a ticket counts
// itself done after its last record access, and the last one of its chain
// returns the chain's ticket, done counter and records to zero.  The buffer
// is zeroed once when it is allocated (tpl_a2b_sync_clear) and stays zero
// between calls.

constexpr int kFwdPay = kF;  // per chunk: aggregate or inclusive frame
constexpr int kBwdPay = 6;   // per chunk: aggregate or inclusive (T, F)

struct ChunkArgs {
    const void* angles;      // [B][3][L]
    void* coords;            // forward output [B][9L]
    const void* grad;        // backward input [B][9L]
    void* grad_angles;       // backward output [B][3][L]
    const int32_t* lengths;  // [B]
    uint32_t* status;
    double* ckpt;            // [B][nk][kF]: entry k is the frame of atom 3kK (N of
                             // residue kK), written for 0 < kK < n; null: none
    uint32_t* tickets;       // [8][sbx]: chain b at (b mod 8) * sbx + b / 8
    uint32_t* done;          // [8][sbx], the same
    uint64_t* pay;           // [B][prec] record words: chunk c of chain b at b * prec + c * PAY
    int sbx;                 // words per stripe of tickets and done (a multiple of 64)
    int prec;                // record words per chain (a multiple of 32)
    int B, L;
    int nc;                  // chunk slots per chain
    int nk, K;               // checkpoints per chain, residues between checkpoints
    int bx;                  // chains per stripe: ceil(B / 8)
};

// A chain's look-back words.  The chains of one stripe run on one XCD, and no
// 256-byte block holds words of two stripes: tickets and done counters are
// stored stripe by stripe in blocks of 64 words, and each chain's records
// start on a 256-byte boundary.  So every cache line of the sync buffer is
// written and read by one XCD only, and no line is shared between the
// non-coherent L2s of two XCDs.
__device__ __forceinline__ uint32_t* ticket_of(const ChunkArgs& a, int b) {
    return a.tickets + (size_t)(b & 7) * a.sbx + (b >> 3);
}
__device__ __forceinline__ uint32_t* done_of(const ChunkArgs& a, int b) {
    return a.done + (size_t)(b & 7) * a.sbx + (b >> 3);
}
__device__ __forceinline__ uint64_t* records_of(const ChunkArgs& a, int b) {
    return a.pay + (size_t)b * a.prec;
}


// Positions.  Stripe x = blockIdx.x mod 8 holds the chains b = x + 8i, i < m
// (m = the stripe's chains) and m*nc positions j; workgroup w takes position
// w / 8 of stripe w mod 8.
//   * Positions j < T, T = sum over the stripe of nreal_i, are the chunks
//     that hold residues, in rows: row r lists, in stripe order, the chains
//     with nreal_i > r (cnt_r of them), and row r's position of a chain
//     takes its r-th chunk in processing order.
//   * Positions T <= j < m*nc are the padding slots, in rows as well: row r
//     lists the chains with nreal_i <= r, whose slot r is past the chain.
// So the work comes first in every stripe, spread over every SIMD by the
// dispatcher, and the zeros after it.  The row counts are popcounts of
// ballots over the stripe's chunk counts (read once per wave, at most
// kMaxStripe of them) and a position's row is found by a scalar walk over
// them, about 10 instructions per row.
// With more than kMaxRows rows or kMaxStripe chains, positions go in slot
// order instead (row r, chain i), the chains of row r rotated by
// r * kRowTurn so that one chain's chunks do not follow each other at the
// period of the dispatcher's placement.
constexpr int kStripeLanes = 8;  // chains per lane
constexpr int kMaxStripe = 64 * kStripeLanes;
constexpr int kMaxRows = 32;
constexpr int kRowTurn = 37;

struct Slot {
    int b;      // chain
    int r;      // row: chunk rank in processing order (real) or padding slot
    int n;      // chain length, clamped to [0, L]
    int nreal;  // chunks that hold residues
    bool real;
    bool bad;   // lengths[b] was outside [0, L]
};

__device__ __forceinline__ Slot make_slot(const ChunkArgs& a, int C, int b, int r) {
    const int raw = a.lengths[b];
    const int n = raw < 0 ? 0 : (raw > a.L ? a.L : raw);
    const int nreal = (n + C - 1) / C;
    return Slot{b, r, n, nreal, r < nreal, raw != n};
}

struct Stripe {
    int x, m;     // stripe and its chain count
    int kk;       // lengths per lane in use; 0: slot order
    int n[kStripeLanes];       // clamped length of chain x + 8 (lane + 64k)
    uint64_t bad[kStripeLanes];  // ballot: that length was outside [0, L]
};

// False if the stripe has no chains.  The stripe's lengths stay in
// registers, so finding a slot needs no further load.
__device__ __forceinline__ bool stripe_init(const ChunkArgs& a, Stripe& st) {
    const int lane = threadIdx.x;
    st.x = (int)(blockIdx.x & 7);
    st.m = a.B > st.x ? (a.B - st.x + 7) / 8 : 0;
    st.kk = (st.m <= kMaxStripe && a.nc <= kMaxRows) ? (st.m + 63) / 64 : 0;
#pragma unroll
    for (int k = 0; k < kStripeLanes; ++k) {
        const int i = lane + 64 * k;
        int n = 0;
        bool bad = false;
        if (k < st.kk && i < st.m) {
            const int raw = a.lengths[st.x + 8 * i];
            n = raw < 0 ? 0 : (raw > a.L ? a.L : raw);
            bad = raw != n;
        }
        st.n[k] = n;
        st.bad[k] = __builtin_amdgcn_ballot_w64(bad);
    }
    return st.m > 0;
}

// The slot at position j < m * nc of the stripe.
__device__ __forceinline__ Slot slot_at(const ChunkArgs& a, int C, const Stripe& st, int j) {
    const int lane = threadIdx.x;
    const int m = st.m, kk = st.kk;
    if (kk == 0) {  // slot order, rows rotated
        const int r = j / m;
        const int i = (j - r * m + (int)(((int64_t)r * kRowTurn) % m)) % m;
        return make_slot(a, C, st.x + 8 * i, r);
    }
    int nr[kStripeLanes];  // chunk counts
#pragma unroll
    for (int k = 0; k < kStripeLanes; ++k) nr[k] = (st.n[k] + C - 1) / C;
    // cnt_r = #{i : nreal_i > r}; real rows first, then padding rows of
    // m - cnt_r slots each.
    int row_start = 0, r = -1, start = 0;
    bool real = true;
    for (int pass = 0; pass < 2 && r < 0; ++pass) {
        for (int rr = 0; rr < a.nc; ++rr) {
            int cnt = 0;
#pragma unroll
            for (int k = 0; k < kStripeLanes; ++k)
                if (k < kk) cnt += __builtin_popcountll(__builtin_amdgcn_ballot_w64(nr[k] > rr));
            const int size = pass == 0 ? cnt : m - cnt;
            if (j < row_start + size) {
                r = rr;
                start = row_start;
                real = pass == 0;
                break;
            }
            row_start += size;
        }
    }
    int q = j - start;  // rank within the row
    int b = st.x, n = 0;
    bool found = false, bad = false;
#pragma unroll
    for (int k = 0; k < kStripeLanes; ++k) {
        if (k >= kk) break;
        const bool in = (lane + 64 * k < m) && (real ? nr[k] > r : nr[k] <= r);
        const uint64_t mask = __builtin_amdgcn_ballot_w64(in);
        const int cnt = __builtin_popcountll(mask);
        if (!found && q < cnt) {
            const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
            const int hl = __builtin_ctzll(__builtin_amdgcn_ballot_w64(in && below == q));
            b = st.x + 8 * (hl + 64 * k);
            n = __builtin_amdgcn_readlane(st.n[k], hl);
            bad = (st.bad[k] >> hl) & 1ull;
            found = true;
        } else if (!found) {
            q -= cnt;
        }
    }
    const int nreal = (n + C - 1) / C;
    return Slot{b, r, n, nreal, r < nreal, bad};
}

// Called once by every wave that took a ticket of chain b, after its last
// access to the chain's records, in two halves so that the counter's round
// trip overlaps the wave's own stores: chain_done_issue counts the wave done
// (after its record stores have landed) and returns the counter's old value,
// chain_done_finish lets the last of the nreal waves return the chain's
// records (PAY words per chunk), ticket and counter to zero: every other wave
// has finished with them, and the next call, of whatever shape, finds the
// buffer as it was before this one (all zero).
__device__ __forceinline__ uint32_t chain_done_issue(const ChunkArgs& a, int b) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);  // this wave's record stores have landed
    uint32_t old = 0;
    if (threadIdx.x == 0) old = atomic_inc(done_of(a, b));
    return old;
}
template <int PAY>
__device__ __forceinline__ void chain_done_finish(const ChunkArgs& a, int b, int nreal, uint32_t old) {
    if (__builtin_amdgcn_readfirstlane(old) == (uint32_t)(nreal - 1)) {
        uint64_t* pay = records_of(a, b);
        for (int j = threadIdx.x; j < nreal * PAY; j += 64) record_store(pay + j, 0ull);
        if (threadIdx.x == 0) {
            word_store(ticket_of(a, b), 0u);
            word_store(done_of(a, b), 0u);
        }
    }
}

// The ticket of a wave of chain b (nreal chunks), or -1 if it is outside
// [0, nreal): the sync buffer was not clean (TPL_A2B_STATUS_SYNC reported;
// the wave then touches nothing of the chain).  In two halves:
// ticket_issue sends the increment (lane 0) and ticket_read waits for its
// result.  Vector-memory results return in issue order, so loads issued
// between the two halves do not delay the ticket.
__device__ __forceinline__ uint32_t ticket_issue(const ChunkArgs& a, int b) {
    uint32_t v = 0;
    if (threadIdx.x == 0) v = atomic_inc(ticket_of(a, b));
    return v;
}
__device__ __forceinline__ int ticket_read(const ChunkArgs& a, uint32_t v, int nreal) {
    const int t = (int)__builtin_amdgcn_readfirstlane(v);
    if (t < 0 || t >= nreal) {
        if (threadIdx.x == 0) set_status(a.status, TPL_A2B_STATUS_SYNC);
        return -1;
    }
    return t;
}
__device__ __forceinline__ int chain_ticket(const ChunkArgs& a, int b, int nreal) {
    return ticket_read(a, ticket_issue(a, b), nreal);
}

// Lane 0 publishes the uniform values v[0..N) as the record at p.
template <int N>
__device__ __forceinline__ void publish(uint64_t* p, const double (&v)[N]) {
    if (threadIdx.x == 0) {
#pragma unroll
        for (int e = 0; e < N; ++e)
            record_store(p + e, static_cast<uint64_t>(__double_as_longlong(v[e])) | 1ull);
    }
}

// Waits until the record at p (read by every valid lane) is written and
// returns it in v (bounded: false after kSpinLimit polls, with
// TPL_A2B_STATUS_TIMEOUT reported).  Invalid lanes read nothing.
//
// The wait is gentle, because polling competes with the waves it waits for.
// Waves of a CU share its vector-memory path, and a wave that is to publish
// a record still has vector-memory instructions to issue (its angle loads,
// its record stores).  With every valid lane re-reading its whole record
// every 64 clocks, the waiting waves of a CU kept a wave there that held an
// earlier chunk from issuing them for milliseconds to seconds (traced on
// MI355X: the holder took its ticket, then reached its next vector-memory
// instruction only when its waiters gave up), and the whole chain waited on
// it.  So: the first poll reads the whole record (usually it is there); later
// polls come only from lanes whose record is still missing, read one word
// (the last, then the whole record once it shows: the words of a record land
// in any order, so each is checked), and sleep 64 * nap clocks between, nap
// doubling up to kMaxNap.
constexpr int kSpinLimit = 1 << 20;
constexpr int kMaxNap = 32;
template <int N>
__device__ __forceinline__ bool wait_record(const uint64_t* p, bool valid, double (&v)[N],
                                            uint32_t* status) {
    bool ok = !valid;
    int nap = 1;
    for (int spins = 0; spins < kSpinLimit; ++spins) {
        if (!ok && (spins == 0 || record_load(p + (N - 1)) != 0)) {
            ok = true;
#pragma unroll
            for (int e = 0; e < N; ++e) v[e] = __longlong_as_double(static_cast<long long>(record_load(p + e)));
#pragma unroll
            for (int e = 0; e < N; ++e) ok = ok && __double_as_longlong(v[e]) != 0;
        }
        if (__builtin_amdgcn_ballot_w64(!ok) == 0) return true;
        for (int z = 0; z < nap; ++z) __builtin_amdgcn_s_sleep(1);
        nap = nap < kMaxNap ? 2 * nap : kMaxNap;
    }
    if (threadIdx.x == 0) set_status(status, TPL_A2B_STATUS_TIMEOUT);
    return false;
}

// Look-back in fixed groups of 64 chunks (0-63, 64-127, ...), so that every
// sum and product is taken in the same order whatever the timing, and the
// results are bitwise reproducible:
//   * forward, chunk c of group g = c / 64:
//       P_c = I(64g - 1) * A_{64g} * ... * A_{c-1},
//     with A_j the aggregate frame of chunk j and I(64g - 1) the inclusive
//     prefix of the previous group's last chunk (g > 0);
//   * backward, chunk c of group g:
//       S_c = A_{c+1} + ... + A_{64g+63} + I(64g + 64),
//     with A_j chunk j's sum and I(64g + 64) the inclusive suffix of the next
//     group's first chunk (where the chain has it).
// Aggregates are published right after a chunk's own scan, so a chunk never
// waits on a chain of look-backs inside its group; only the group-boundary
// chunks publish inclusive values (in the record their aggregate would use:
// no chunk reads a boundary's aggregate), and those form one chain per 64
// chunks.  Lane k of the window reads one chunk's aggregate; the window's
// product or sum is one wave scan in lane order (identity or zero in unused
// lanes), skipped when the window is one chunk.

// Forward: P <- the exclusive prefix of chunk c >= 1 of chain b (P enters
// as I; uniform over the wave).
__device__ void fwd_lookback(const ChunkArgs& a, int b, int c, double P[kF]) {
    const int lane = threadIdx.x;
    const uint64_t* pay = records_of(a, b);
    const int lo = c & ~63;
    if (lo > 0) {  // the previous group's inclusive prefix, read by lane 0
        double I[kF] = {};
        if (!wait_record<kF>(pay + (size_t)(lo - 1) * kFwdPay, lane == 0, I, a.status)) return;
#pragma unroll
        for (int e = 0; e < kF; ++e) P[e] = read_lane(I[e], 0);
    }
    if (c == lo) return;
    const int j = lo + lane;
    const bool valid = j < c;
    double v[kF] = {};
    if (!wait_record<kF>(pay + (size_t)(valid ? j : lo) * kFwdPay, valid, v, a.status)) return;
    if (!valid) {
#pragma unroll
        for (int e = 0; e < kF; ++e) v[e] = ident(e);
    }
    if (c - lo > 1) wave_scan_frames(v);  // lane 63: A_lo * ... * A_{c-1}
    const int src = c - lo > 1 ? 63 : 0;  // a window of one chunk: lane 0
    double W[kF];
#pragma unroll
    for (int e = 0; e < kF; ++e) W[e] = read_lane(v[e], src);
    frame_rmul_to(P, W, v);  // v = P * W
#pragma unroll
    for (int e = 0; e < kF; ++e) P[e] = v[e];
}

// Backward: carry <- the sum of (T, F) over the chunks after chunk c of
// chain b (nreal chunks); carry enters as 0.
__device__ void bwd_lookback(const ChunkArgs& a, int b, int c, int nreal, double carry[6]) {
    const int lane = threadIdx.x;
    const uint64_t* pay = records_of(a, b);
    const int lo = c & ~63, hi = min(lo + 64, nreal);  // c's group, clipped to the chain
    const int j = c + 1 + lane;  // lane k reads chunk c + 1 + k
    const bool valid = j < hi;
    double v[6] = {};  // zero in unused lanes
    if (!wait_record<6>(pay + (size_t)(valid ? j : c) * kBwdPay, valid, v, a.status)) return;
    if (hi - c - 1 > 1) wave_scan_sum<6>(v);  // lane 63: the window's sum
    const int src = hi - c - 1 > 1 ? 63 : 0;   // a window of one chunk: lane 0
#pragma unroll
    for (int e = 0; e < 6; ++e) carry[e] = read_lane(v[e], src);
    if (hi < nreal) {  // the next group's inclusive suffix, read by lane 0
        double in[6] = {};
        if (!wait_record<6>(pay + (size_t)hi * kBwdPay, lane == 0, in, a.status)) return;
#pragma unroll
        for (int e = 0; e < 6; ++e) carry[e] += read_lane(in[e], 0);
    }
}

// ------------------------------------------------------------ forward chunk
// Steps of one chunk (chunk c of chain b, first residue r0 = c*C):
//   1. the lanes load their R angles per row straight into registers (zero
//      past the chain); a slot guesses c = s and loads before its ticket
//      returns, reloading in the rare case the guess was wrong;
//   2. pass 1: per lane, FP64 sincos and the serial product of its R
//      residues; positions of the lane's atoms 1 .. 3R-1 in the lane's own
//      frame go to the lane's LDS slot as FP32 (|p| <= 3R * 1.53 A);
//   3. DPP scan of the lane frames; look-back over earlier chunks (P);
//   4. E = P * (lane's prefix in the wave) is the frame of the lane's first
//      atom; lanes on a backward-chunk boundary save it as a checkpoint;
//   5. pass 2: atom 0 is the translation of E, atom a >= 1 is E p_a, written
//      over p_a in LDS;
//   6. the wave stores its 9*C-float span with 16-byte stores from LDS,
//      zeros past the chain.
// One slot of the forward; s_tab is filled on the wave's first chunk (tab).
template <typename T, int R>
__device__ __forceinline__ void fwd_slot(const ChunkArgs& a, const Slot& sl, T* s_x, double* s_tab,
                                         bool& tab) {
    constexpr int C = 64 * R;
    constexpr int E9 = 9 * R;  // output floats per lane = the lane's LDS stride
    T* s_l = s_x + threadIdx.x * E9;
    const int lane = threadIdx.x;
    const int L = a.L, b = sl.b, n = sl.n, nreal = sl.nreal, s = sl.r;
    if (sl.bad && s == 0 && lane == 0) set_status(a.status, TPL_A2B_STATUS_LENGTH);
    T* xrow = static_cast<T*>(a.coords) + (size_t)b * 9 * L;
    if (!sl.real) {  // past the chain: the first such slot writes all its zeros
        if (s == nreal) {
            const int f0 = 9 * s * C;
            wave_store_run<T>(xrow + f0, 9 * L - f0, 0, s_x, lane);
        }
        return;
    }
    const bool fill = !tab;
    if (fill) {
        for (int i = lane; i < 128; i += 64) s_tab[i] = kTabCS[i];
    }
    const T* arow = static_cast<const T*>(a.angles) + (size_t)b * 3 * L;
    T ang[3][R];
    {
        const int rl = s * C + lane * R;
        lane_load_rows<T, R>(arow + rl, L, min(max(n - rl, 0), R), L - rl, ang);
    }
    const int c = nreal > 1 ? chain_ticket(a, b, nreal) : 0;
    if (c < 0) return;
    const int r0 = c * C;
    const int rl = r0 + lane * R;
    const int cnt = min(max(n - rl, 0), R);  // residues of the chain in this lane
    if (c != s) lane_load_rows<T, R>(arow + rl, L, cnt, L - rl, ang);
    if (fill) {
        __syncthreads();  // s_tab
        tab = true;
    }

    // Pass 1.
    double E[kF];
    frame_identity(E);
#pragma unroll
    for (int q = 0; q < R; ++q) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double sn, cs;
            sincos_tab((double)ang[k][q], s_tab, &sn, &cs);
            if (k == 0) frame_step<0>(E, cs, sn);
            if (k == 1) frame_step<1>(E, cs, sn);
            if (k == 2) frame_step<2>(E, cs, sn);
            const int at = 3 * q + k + 1;  // lane atom placed by this step
            if (at < 3 * R) {
                s_l[3 * at] = (T)E[9];
                s_l[3 * at + 1] = (T)E[10];
                s_l[3 * at + 2] = (T)E[11];
            }
        }
    }
    wave_scan_frames(E);
    double agg[kF];  // the chunk's aggregate (uniform)
#pragma unroll
    for (int e = 0; e < kF; ++e) agg[e] = read_lane(E[e], 63);
    wave_shift_frames(E);

    // Chunks of one chain: a chunk's aggregate is read by the later chunks
    // of its group; the last chunk of a group publishes its inclusive prefix
    // instead, for the next group; the chain's last chunk publishes nothing.
    if (nreal > 1) {
        uint64_t* my = records_of(a, b) + (size_t)c * kFwdPay;
        const bool last = c == nreal - 1;
        const bool boundary = (c & 63) == 63;
        if (!last && !boundary) publish<kF>(my, agg);
        if (c > 0) {
            double P[kF];
            frame_identity(P);
            fwd_lookback(a, b, c, P);
            frame_lmul(P, E);
            if (!last && boundary) {
                double incl[kF];
                frame_rmul_to(P, agg, incl);
                publish<kF>(my, incl);
            }
        }
    }

    // Checkpoint at a backward-chunk boundary inside the chain.
    if (a.ckpt && rl > 0 && rl < n && (rl & (a.K - 1)) == 0) {
        double* d = a.ckpt + ((size_t)b * a.nk + rl / a.K) * kF;
#pragma unroll
        for (int e = 0; e < kF; ++e) d[e] = E[e];
    }

    // Pass 2, in place in the lane's slot.
    if (cnt > 0) {
        // Sum of the lane's coordinates: NaN or inf anywhere shows.
        // Coordinates past the chain come from zero angles: finite.
        float chk = 0.f;
#pragma unroll
        for (int at = 0; at < 3 * R; ++at) {
            double x[3];
            if (at == 0) {
                x[0] = E[9];
                x[1] = E[10];
                x[2] = E[11];
            } else {
                affine(E, (double)s_l[3 * at], (double)s_l[3 * at + 1], (double)s_l[3 * at + 2], x);
            }
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                const T o = (T)x[cc];
                s_l[3 * at + cc] = o;
                chk += (float)o;
            }
        }
        if (!__builtin_isfinite(chk)) set_status(a.status, TPL_A2B_STATUS_NAN);
    }
    const uint32_t done_old = nreal > 1 ? chain_done_issue(a, b) : 0u;
    wave_lds_sync();
    const int g0 = 9 * r0;
    const int wcount = min(9 * C, 9 * L - g0);
    wave_store_run<T>(xrow + g0, wcount, min(max(9 * n - g0, 0), wcount), s_x, lane);
    if (nreal > 1) chain_done_finish<kFwdPay>(a, b, nreal, done_old);
}

template <typename T, int R>
__global__ void __launch_bounds__(64) a2b_fwd_chunk(ChunkArgs a) {
    constexpr int C = 64 * R;
    extern __shared__ __align__(16) unsigned char s_dyn[];
    __shared__ double s_tab[128];
    T* s_x = reinterpret_cast<T*>(s_dyn);
    Stripe st;
    const int j = (int)(blockIdx.x >> 3);
    if (!stripe_init(a, st) || j >= st.m * a.nc) return;
    bool tab = false;
    fwd_slot<T, R>(a, slot_at(a, C, st, j), s_x, s_tab, tab);
}

// ------------------------------------------------------------ forward chain
// Chains of at most kChainMax chunks (L <= kChainMax * C) can skip the
// look-back: one workgroup per chain, one wave per chunk slot, and the
// chunks' aggregates meet in LDS across one barrier.  No tickets, records or
// done counters, and each wave finds its chunk from the workgroup index.
// Wave w takes slot (w + b) mod W, so that a chain's first chunk, the only
// one every chain has, does not always fall on the same wave (and SIMD).
// Otherwise as fwd_slot: pass 1, DPP scan, E = (earlier chunks' aggregates
// in order) * (lane prefix), checkpoints, pass 2, stores; the slot right
// past the chain writes its zeros.
constexpr int kChainMax = 4;

template <typename T, int R, int W>
__global__ void __launch_bounds__(64 * W) a2b_fwd_chain(ChunkArgs a) {
    constexpr int C = 64 * R;
    constexpr int E9 = 9 * R;
    extern __shared__ __align__(16) unsigned char s_dyn[];
    __shared__ double s_tab[128];
    __shared__ double s_agg[W][kF];
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const int b = (int)blockIdx.x, L = a.L;
    const int raw = a.lengths[b];
    const int n = raw < 0 ? 0 : (raw > L ? L : raw);
    const int nreal = (n + C - 1) / C;
    const int c = (wave + b) % W;  // this wave's chunk slot
    const bool real = c < nreal;
    if (raw != n && threadIdx.x == 0) set_status(a.status, TPL_A2B_STATUS_LENGTH);
    for (int i = (int)threadIdx.x; i < 128; i += 64 * W) s_tab[i] = kTabCS[i];
    T* s_x = reinterpret_cast<T*>(s_dyn) + (size_t)wave * 9 * C;
    T* s_l = s_x + lane * E9;
    T* xrow = static_cast<T*>(a.coords) + (size_t)b * 9 * L;
    const T* arow = static_cast<const T*>(a.angles) + (size_t)b * 3 * L;
    const int r0 = c * C;
    const int rl = r0 + lane * R;
    const int cnt = min(max(n - rl, 0), R);
    T ang[3][R];
    if (real) lane_load_rows<T, R>(arow + rl, L, cnt, L - rl, ang);
    __syncthreads();  // s_tab

    // Pass 1 and the wave's scan.
    double E[kF];
    frame_identity(E);
    if (real) {
#pragma unroll
        for (int q = 0; q < R; ++q) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double sn, cs;
                sincos_tab((double)ang[k][q], s_tab, &sn, &cs);
                if (k == 0) frame_step<0>(E, cs, sn);
                if (k == 1) frame_step<1>(E, cs, sn);
                if (k == 2) frame_step<2>(E, cs, sn);
                const int at = 3 * q + k + 1;
                if (at < 3 * R) {
                    s_l[3 * at] = (T)E[9];
                    s_l[3 * at + 1] = (T)E[10];
                    s_l[3 * at + 2] = (T)E[11];
                }
            }
        }
        wave_scan_frames(E);
        double agg[kF];
#pragma unroll
        for (int e = 0; e < kF; ++e) agg[e] = read_lane(E[e], 63);
        if (lane == 0) {
#pragma unroll
            for (int e = 0; e < kF; ++e) s_agg[c][e] = agg[e];
        }
        wave_shift_frames(E);
    }
    __syncthreads();  // the aggregates
    if (real && c > 0) {
        double P[kF], Q[kF];
#pragma unroll
        for (int e = 0; e < kF; ++e) P[e] = s_agg[0][e];
        for (int j = 1; j < c; ++j) {
            frame_rmul_to(P, s_agg[j], Q);
#pragma unroll
            for (int e = 0; e < kF; ++e) P[e] = Q[e];
        }
        frame_lmul(P, E);
    }
    if (!real) {
        if (c == nreal) {  // the first slot past the chain writes all its zeros
            const int f0 = 9 * c * C;
            wave_store_run<T>(xrow + f0, 9 * L - f0, 0, s_x, lane);
        }
        return;
    }

    // Checkpoint at a backward-chunk boundary inside the chain.
    if (a.ckpt && rl > 0 && rl < n && (rl & (a.K - 1)) == 0) {
        double* d = a.ckpt + ((size_t)b * a.nk + rl / a.K) * kF;
#pragma unroll
        for (int e = 0; e < kF; ++e) d[e] = E[e];
    }

    // Pass 2, in place in the lane's slot.
    if (cnt > 0) {
        float chk = 0.f;
#pragma unroll
        for (int at = 0; at < 3 * R; ++at) {
            double x[3];
            if (at == 0) {
                x[0] = E[9];
                x[1] = E[10];
                x[2] = E[11];
            } else {
                affine(E, (double)s_l[3 * at], (double)s_l[3 * at + 1], (double)s_l[3 * at + 2], x);
            }
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                const T o = (T)x[cc];
                s_l[3 * at + cc] = o;
                chk += (float)o;
            }
        }
        if (!__builtin_isfinite(chk)) set_status(a.status, TPL_A2B_STATUS_NAN);
    }
    wave_lds_sync();
    const int g0 = 9 * r0;
    const int wcount = min(9 * C, 9 * L - g0);
    wave_store_run<T>(xrow + g0, wcount, min(max(9 * n - g0, 0), wcount), s_x, lane);
}

// ----------------------------------------------------------- backward chunk
// Chain atom j >= 1 (atoms N_0, CA_0, C_0, N_1, ...) is placed by the
// transform whose angle theta_j is phi_r, psi_r, omega_r for j = 3r+1, 3r+2,
// 3r+3.  theta_j turns every later atom about the bond axis u_j (column 0 of
// A_j) through x_j, so
//   dL/dtheta_j = u_j . sum_{i>=j} (x_i - x_j) x g_i = u_j . (T_j - x_j x F_j),
//   T_j = sum_{i>=j} x_i x g_i,  F_j = sum_{i>=j} g_i
// (the i = j term is zero).  Steps of one chunk (C = 64R = the checkpoint
// stride K):
//   1. the lanes load their angles and the gradients of their 3R atoms;
//   2. pass 1: per lane, sincos and the serial product of its residues from
//      the identity, as in the forward; the lane keeps (cos, sin) of its 3R
//      angles and sums, over its atoms a in the lane's own frame,
//      M = sum p_a g_a^T and F = sum g_a;
//   3. DPP scan of the lane frames; E = checkpoint * (lane's prefix) is the
//      frame of the lane's first atom, and its sum over the lane's atoms is
//      W = (eps : (R_E M) + t_E x F, F);
//   4. suffix sums of W over later lanes (DPP) and later chunks (look-back);
//   5. pass B: from E the lane steps through its 3R transforms again with the
//      kept (cos, sin); at atom j it has x_j and u_j and the suffix
//      S_j = sum over atoms i >= j, gives the gradient, and drops atom j's
//      term from S.  psi and omega of the last residue are 0.
// One slot of the backward; s_tab is filled on the wave's first chunk (tab).
// s_cs [6R][64]: the lanes' (cos, sin) from pass 1 for pass B, element-major
// so that a wave's access is 64 consecutive doubles.
template <typename T, int R>
__device__ __forceinline__ void bwd_slot(const ChunkArgs& a, const Slot& sl, double* s_tab, double* s_cs,
                                         bool& tab) {
    constexpr int C = 64 * R;
    const int lane = threadIdx.x;
    const int L = a.L, b = sl.b, n = sl.n, nreal = sl.nreal, s = sl.r;
    if (sl.bad && s == 0 && lane == 0) set_status(a.status, TPL_A2B_STATUS_LENGTH);
    T* grow = static_cast<T*>(a.grad_angles) + (size_t)b * 3 * L;
    if (!sl.real) {  // past the chain: the first such slot writes all its zeros
        if (s == nreal) {
#pragma unroll
            for (int k = 0; k < 3; ++k)
                wave_store_run<T>(grow + (size_t)k * L + s * C, L - s * C, 0, static_cast<const T*>(nullptr), lane);
        }
        return;
    }
    const bool fill = !tab;
    if (fill) {
        for (int i = lane; i < 128; i += 64) s_tab[i] = kTabCS[i];
    }
    const T* arow = static_cast<const T*>(a.angles) + (size_t)b * 3 * L;
    const T* gin = static_cast<const T*>(a.grad) + (size_t)b * 9 * L;
    T ang[3][R];
    T gr[9 * R];
    // A lane wholly past the chain reads nothing: whole-run reads there
    // would fetch the padding of every chain's last chunk, about half a chunk
    // of gradients and angles per chain (+45% over the valid inputs at
    // B=1024 with lengths 30..1000).
    auto load = [&](int cc) {
        const int r = cc * C + lane * R;
        const int k = min(max(n - r, 0), R);
        if (k > 0) {
            lane_load_rows<T, R>(arow + r, L, k, L - r, ang);
            lane_load<T, 9 * R>(gin + (size_t)9 * r, 9 * k, gr, 9 * (L - r));
        } else {
#pragma unroll
            for (int kk = 0; kk < 3; ++kk)
#pragma unroll
                for (int i = 0; i < R; ++i) ang[kk][i] = (T)0;
#pragma unroll
            for (int i = 0; i < 9 * R; ++i) gr[i] = (T)0;
        }
    };
    double P[kF];  // checkpoint: frame of the chunk's first atom (uniform)
    auto load_ckpt = [&](int cc) {
        if (cc == 0) {
            frame_identity(P);
        } else {
            const double* ck = a.ckpt + ((size_t)b * a.nk + cc) * kF;
#pragma unroll
            for (int e = 0; e < kF; ++e) P[e] = ck[e];
        }
    };
    // The ticket goes out first and the guessed chunk's loads after it, so
    // the ticket returns first and a wrong guess is reloaded at once.
    const int guess = nreal - 1 - s;  // the chunk of this slot if tickets follow slot order
    uint32_t tv = nreal > 1 ? ticket_issue(a, b) : 0u;
    load(guess);
    load_ckpt(guess);
    // Keeps the compiler from reading the ticket (and so waiting for it)
    // before the loads above are issued.
    __asm__ volatile("" : "+v"(tv) : : "memory");
    const int t = nreal > 1 ? ticket_read(a, tv, nreal) : 0;
    if (t < 0) return;
    const int c = nreal - 1 - t;
    if (c != guess) {
        load(c);
        load_ckpt(c);
    }
    const int r0 = c * C;
    const int rl = r0 + lane * R;
    const int cnt = min(max(n - rl, 0), R);
    if (fill) {
        __syncthreads();  // s_tab
        tab = true;
    }

    // Pass 1 in the lane's own frame.
    double f[kF];
    frame_identity(f);
    double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // M[3m + k] = sum_a p_a[m] g_a[k]
    double F[3] = {(double)gr[0], (double)gr[1], (double)gr[2]};  // atom 0 sits at p = 0
#pragma unroll
    for (int q = 0; q < R; ++q) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double sn, cs;
            sincos_tab((double)ang[k][q], s_tab, &sn, &cs);
            s_cs[(6 * q + 2 * k) * 64 + lane] = cs;
            s_cs[(6 * q + 2 * k + 1) * 64 + lane] = sn;
            if (k == 0) frame_step<0>(f, cs, sn);
            if (k == 1) frame_step<1>(f, cs, sn);
            if (k == 2) frame_step<2>(f, cs, sn);
            const int at = 3 * q + k + 1;
            if (at < 3 * R) {
                const double g0 = (double)gr[3 * at], g1 = (double)gr[3 * at + 1],
                             g2 = (double)gr[3 * at + 2];
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    M[3 * m] = fma(f[9 + m], g0, M[3 * m]);
                    M[3 * m + 1] = fma(f[9 + m], g1, M[3 * m + 1]);
                    M[3 * m + 2] = fma(f[9 + m], g2, M[3 * m + 2]);
                }
                F[0] += g0;
                F[1] += g1;
                F[2] += g2;
            }
        }
    }
    wave_scan_frames(f);
    wave_shift_frames(f);
    frame_lmul(P, f);  // f = E, the frame of the lane's first atom

    // W = sum over the lane's atoms of (x x g, g), x = R_E p + t_E:
    // sum (R_E p) x g = eps : (R_E M), with (R_E M)[j][k] = sum_m R[j][m] M[m][k].
    double W[6];
    {
        auto rm = [&](int j, int k) {
            return f[j] * M[k] + f[3 + j] * M[3 + k] + f[6 + j] * M[6 + k];
        };
        W[0] = rm(1, 2) - rm(2, 1) + (f[10] * F[2] - f[11] * F[1]);
        W[1] = rm(2, 0) - rm(0, 2) + (f[11] * F[0] - f[9] * F[2]);
        W[2] = rm(0, 1) - rm(1, 0) + (f[9] * F[1] - f[10] * F[0]);
        W[3] = F[0];
        W[4] = F[1];
        W[5] = F[2];
    }
    // S = sum over later lanes of the chunk, then over later chunks.
    double S[6];
#pragma unroll
    for (int e = 0; e < 6; ++e) S[e] = W[e];
    wave_scan_sum<6>(S);  // inclusive prefix over lanes
    double tot[6];
#pragma unroll
    for (int e = 0; e < 6; ++e) {
        tot[e] = read_lane(S[e], 63);
        S[e] = tot[e] - S[e];
    }
    // Chunks of one chain: a chunk's sum is read by the earlier chunks of
    // its group; the first chunk of a group (after the first group) publishes
    // its inclusive suffix instead, for the previous group; chunk 0 publishes
    // nothing.
    if (nreal > 1) {
        uint64_t* my = records_of(a, b) + (size_t)c * kBwdPay;
        const bool boundary = (c & 63) == 0;
        if (!boundary) publish<6>(my, tot);
        double carry[6] = {0, 0, 0, 0, 0, 0};
        if (c < nreal - 1) bwd_lookback(a, b, c, nreal, carry);
        if (boundary && c > 0) {
            double incl[6];
#pragma unroll
            for (int e = 0; e < 6; ++e) incl[e] = carry[e] + tot[e];
            publish<6>(my, incl);
        }
#pragma unroll
        for (int e = 0; e < 6; ++e) S[e] += carry[e];
    }

    const uint32_t done_old = nreal > 1 ? chain_done_issue(a, b) : 0u;
    if (cnt > 0) {
        // S becomes the suffix from the lane's atom 1: add the lane's own
        // atoms, drop atom 0.
        {
            const double g0 = (double)gr[0], g1 = (double)gr[1], g2 = (double)gr[2];
            S[0] += W[0] - (f[10] * g2 - f[11] * g1);
            S[1] += W[1] - (f[11] * g0 - f[9] * g2);
            S[2] += W[2] - (f[9] * g1 - f[10] * g0);
            S[3] += W[3] - g0;
            S[4] += W[4] - g1;
            S[5] += W[5] - g2;
        }
        T out[3][R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double cs = s_cs[(6 * q + 2 * k) * 64 + lane];
                const double sn = s_cs[(6 * q + 2 * k + 1) * 64 + lane];
                if (k == 0) frame_step<0>(f, cs, sn);
                if (k == 1) frame_step<1>(f, cs, sn);
                if (k == 2) frame_step<2>(f, cs, sn);
                const int j = 3 * q + k + 1;  // lane atom j: x = t(f), u = column 0 of f
                const double x0 = f[9], x1 = f[10], x2 = f[11];
                const double v = f[0] * (S[0] - (x1 * S[5] - x2 * S[4])) +
                                 f[1] * (S[1] - (x2 * S[3] - x0 * S[5])) +
                                 f[2] * (S[2] - (x0 * S[4] - x1 * S[3]));
                const int r = rl + q;
                const bool zero = (r >= n) || (r == n - 1 && k > 0);
                out[k][q] = zero ? (T)0 : (T)v;
                if (j < 3 * R) {
                    const double g0 = (double)gr[3 * j], g1 = (double)gr[3 * j + 1],
                                 g2 = (double)gr[3 * j + 2];
                    S[0] -= x1 * g2 - x2 * g1;
                    S[1] -= x2 * g0 - x0 * g2;
                    S[2] -= x0 * g1 - x1 * g0;
                    S[3] -= g0;
                    S[4] -= g1;
                    S[5] -= g2;
                }
            }
        }
        bool bad = false;
        const int crow = min(max(L - rl, 0), R);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
#pragma unroll
            for (int q = 0; q < R; ++q) bad |= (out[k][q] != out[k][q]);
            lane_store<T, R>(grow + (size_t)k * L + rl, crow, out[k]);
        }
        if (bad) set_status(a.status, TPL_A2B_STATUS_NAN);
    } else {
        const int crow = min(max(L - rl, 0), R);
        if (crow > 0) {
#pragma unroll
            for (int k = 0; k < 3; ++k) lane_store_zero<T, R>(grow + (size_t)k * L + rl, crow);
        }
    }
    if (nreal > 1) chain_done_finish<kBwdPay>(a, b, nreal, done_old);
}

template <typename T, int R>
__global__ void __launch_bounds__(64) a2b_bwd_chunk(ChunkArgs a) {
    constexpr int C = 64 * R;
    extern __shared__ double s_cs[];  // [6R][64]
    __shared__ double s_tab[128];
    Stripe st;
    const int j = (int)(blockIdx.x >> 3);
    if (!stripe_init(a, st) || j >= st.m * a.nc) return;
    bool tab = false;
    bwd_slot<T, R>(a, slot_at(a, C, st, j), s_tab, s_cs, tab);
}

// ----------------------------------------------------------- backward chain
// The backward of chains of at most kChainMax chunks: one workgroup per
// chain, one wave per chunk slot (slot (w + b) mod W, as in the forward), and
// each chunk's sum over its atoms meets the others in LDS across one barrier:
// chunk c adds the sums of chunks c+1 .. nreal-1, in that order.  Otherwise
// as bwd_slot.
template <typename T, int R, int W>
__global__ void __launch_bounds__(64 * W) a2b_bwd_chain(ChunkArgs a) {
    constexpr int C = 64 * R;
    extern __shared__ double s_dynd[];  // [W][6R][64]: each wave's (cos, sin)
    __shared__ double s_tab[128];
    __shared__ double s_tot[W][6];
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const int b = (int)blockIdx.x, L = a.L;
    const int raw = a.lengths[b];
    const int n = raw < 0 ? 0 : (raw > L ? L : raw);
    const int nreal = (n + C - 1) / C;
    const int c = (wave + b) % W;  // this wave's chunk slot
    const bool real = c < nreal;
    if (raw != n && threadIdx.x == 0) set_status(a.status, TPL_A2B_STATUS_LENGTH);
    for (int i = (int)threadIdx.x; i < 128; i += 64 * W) s_tab[i] = kTabCS[i];
    double* s_cs = s_dynd + (size_t)wave * 6 * R * 64;
    T* grow = static_cast<T*>(a.grad_angles) + (size_t)b * 3 * L;
    const T* arow = static_cast<const T*>(a.angles) + (size_t)b * 3 * L;
    const T* gin = static_cast<const T*>(a.grad) + (size_t)b * 9 * L;
    const int r0 = c * C;
    const int rl = r0 + lane * R;
    const int cnt = min(max(n - rl, 0), R);
    T ang[3][R];
    T gr[9 * R];
    double P[kF];  // checkpoint: frame of the chunk's first atom (uniform)
    if (real) {
        if (cnt > 0) {  // a lane wholly past the chain reads nothing
            lane_load_rows<T, R>(arow + rl, L, cnt, L - rl, ang);
            lane_load<T, 9 * R>(gin + (size_t)9 * rl, 9 * cnt, gr, 9 * (L - rl));
        } else {
#pragma unroll
            for (int kk = 0; kk < 3; ++kk)
#pragma unroll
                for (int i = 0; i < R; ++i) ang[kk][i] = (T)0;
#pragma unroll
            for (int i = 0; i < 9 * R; ++i) gr[i] = (T)0;
        }
        if (c == 0) {
            frame_identity(P);
        } else {
            const double* ck = a.ckpt + ((size_t)b * a.nk + c) * kF;
#pragma unroll
            for (int e = 0; e < kF; ++e) P[e] = ck[e];
        }
    }
    __syncthreads();  // s_tab

    double f[kF], W6[6], S[6];
    if (real) {
        // Pass 1 in the lane's own frame.
        frame_identity(f);
        double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double F[3] = {(double)gr[0], (double)gr[1], (double)gr[2]};
#pragma unroll
        for (int q = 0; q < R; ++q) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double sn, cs;
                sincos_tab((double)ang[k][q], s_tab, &sn, &cs);
                s_cs[(6 * q + 2 * k) * 64 + lane] = cs;
                s_cs[(6 * q + 2 * k + 1) * 64 + lane] = sn;
                if (k == 0) frame_step<0>(f, cs, sn);
                if (k == 1) frame_step<1>(f, cs, sn);
                if (k == 2) frame_step<2>(f, cs, sn);
                const int at = 3 * q + k + 1;
                if (at < 3 * R) {
                    const double g0 = (double)gr[3 * at], g1 = (double)gr[3 * at + 1],
                                 g2 = (double)gr[3 * at + 2];
#pragma unroll
                    for (int m = 0; m < 3; ++m) {
                        M[3 * m] = fma(f[9 + m], g0, M[3 * m]);
                        M[3 * m + 1] = fma(f[9 + m], g1, M[3 * m + 1]);
                        M[3 * m + 2] = fma(f[9 + m], g2, M[3 * m + 2]);
                    }
                    F[0] += g0;
                    F[1] += g1;
                    F[2] += g2;
                }
            }
        }
        wave_scan_frames(f);
        wave_shift_frames(f);
        frame_lmul(P, f);  // f = E, the frame of the lane's first atom
        auto rm = [&](int j, int k) {
            return f[j] * M[k] + f[3 + j] * M[3 + k] + f[6 + j] * M[6 + k];
        };
        W6[0] = rm(1, 2) - rm(2, 1) + (f[10] * F[2] - f[11] * F[1]);
        W6[1] = rm(2, 0) - rm(0, 2) + (f[11] * F[0] - f[9] * F[2]);
        W6[2] = rm(0, 1) - rm(1, 0) + (f[9] * F[1] - f[10] * F[0]);
        W6[3] = F[0];
        W6[4] = F[1];
        W6[5] = F[2];
#pragma unroll
        for (int e = 0; e < 6; ++e) S[e] = W6[e];
        wave_scan_sum<6>(S);  // inclusive prefix over lanes
        double tot[6];
#pragma unroll
        for (int e = 0; e < 6; ++e) {
            tot[e] = read_lane(S[e], 63);
            S[e] = tot[e] - S[e];
        }
        if (lane == 0) {
#pragma unroll
            for (int e = 0; e < 6; ++e) s_tot[c][e] = tot[e];
        }
    }
    __syncthreads();  // the chunks' sums
    if (!real) {
        if (c == nreal) {  // the first slot past the chain writes all its zeros
#pragma unroll
            for (int k = 0; k < 3; ++k)
                wave_store_run<T>(grow + (size_t)k * L + c * C, L - c * C, 0, static_cast<const T*>(nullptr), lane);
        }
        return;
    }
    for (int j = c + 1; j < nreal; ++j) {
#pragma unroll
        for (int e = 0; e < 6; ++e) S[e] += s_tot[j][e];
    }
    if (cnt > 0) {
        // S becomes the suffix from the lane's atom 1: add the lane's own
        // atoms, drop atom 0.
        {
            const double g0 = (double)gr[0], g1 = (double)gr[1], g2 = (double)gr[2];
            S[0] += W6[0] - (f[10] * g2 - f[11] * g1);
            S[1] += W6[1] - (f[11] * g0 - f[9] * g2);
            S[2] += W6[2] - (f[9] * g1 - f[10] * g0);
            S[3] += W6[3] - g0;
            S[4] += W6[4] - g1;
            S[5] += W6[5] - g2;
        }
        T out[3][R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double cs = s_cs[(6 * q + 2 * k) * 64 + lane];
                const double sn = s_cs[(6 * q + 2 * k + 1) * 64 + lane];
                if (k == 0) frame_step<0>(f, cs, sn);
                if (k == 1) frame_step<1>(f, cs, sn);
                if (k == 2) frame_step<2>(f, cs, sn);
                const int j = 3 * q + k + 1;
                const double x0 = f[9], x1 = f[10], x2 = f[11];
                const double v = f[0] * (S[0] - (x1 * S[5] - x2 * S[4])) +
                                 f[1] * (S[1] - (x2 * S[3] - x0 * S[5])) +
                                 f[2] * (S[2] - (x0 * S[4] - x1 * S[3]));
                const int r = rl + q;
                const bool zero = (r >= n) || (r == n - 1 && k > 0);
                out[k][q] = zero ? (T)0 : (T)v;
                if (j < 3 * R) {
                    const double g0 = (double)gr[3 * j], g1 = (double)gr[3 * j + 1],
                                 g2 = (double)gr[3 * j + 2];
                    S[0] -= x1 * g2 - x2 * g1;
                    S[1] -= x2 * g0 - x0 * g2;
                    S[2] -= x0 * g1 - x1 * g0;
                    S[3] -= g0;
                    S[4] -= g1;
                    S[5] -= g2;
                }
            }
        }
        bool bad = false;
        const int crow = min(max(L - rl, 0), R);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
#pragma unroll
            for (int q = 0; q < R; ++q) bad |= (out[k][q] != out[k][q]);
            lane_store<T, R>(grow + (size_t)k * L + rl, crow, out[k]);
        }
        if (bad) set_status(a.status, TPL_A2B_STATUS_NAN);
    } else {
        const int crow = min(max(L - rl, 0), R);
        if (crow > 0) {
#pragma unroll
            for (int k = 0; k < 3; ++k) lane_store_zero<T, R>(grow + (size_t)k * L + rl, crow);
        }
    }
}

__global__ void a2b_zero_words(uint32_t* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// ------------------------------------------------------------- plain form
// Reference algorithm as straightforward HIP, computing in T.
template <typename T>
struct PlainFrame {
    T v[kF];
};

template <typename T>
__device__ __forceinline__ void plain_compose(const T A[kF], const T B[kF], T out[kF]) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            T v = A[i] * B[3 * j] + A[3 + i] * B[3 * j + 1] + A[6 + i] * B[3 * j + 2];
            out[3 * j + i] = (j == 3) ? v + A[9 + i] : v;
        }
}

template <typename T>
__device__ __forceinline__ void plain_sincos(T x, T* s, T* c);
template <>
__device__ __forceinline__ void plain_sincos<float>(float x, float* s, float* c) {
    sincosf(x, s, c);
}
template <>
__device__ __forceinline__ void plain_sincos<double>(double x, double* s, double* c) {
    sincos(x, s, c);
}

// B(theta, beta_k, d_k) as a column-stored 3x4 (Math/cMatrix44.cpp:47-52).
template <typename T>
__device__ __forceinline__ void plain_dihedral(int k, T th, T out[kF]) {
    const Bond bd = bond(k);
    T s, c;
    plain_sincos<T>(th, &s, &c);
    const T cb = (T)bd.cb, sb = (T)bd.sb, d = (T)bd.d;
    out[0] = cb;  out[1] = 0;  out[2] = -sb;          // column 0
    out[3] = s * sb; out[4] = c; out[5] = s * cb;      // column 1
    out[6] = c * sb; out[7] = -s; out[8] = c * cb;     // column 2
    out[9] = d * cb; out[10] = 0; out[11] = -d * sb;   // translation
}

template <typename T>
__global__ void a2b_fwd_plain(const T* angles, T* coords, const int32_t* lengths,
                              T* frames, uint32_t* status, int B, int L) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    int n = lengths[b];
    if (n < 0 || n > L) {
        set_status(status, TPL_A2B_STATUS_LENGTH);
        n = n < 0 ? 0 : L;
    }
    const T* ang = angles + (size_t)b * 3 * L;
    T* x = coords + (size_t)b * 9 * L;
    T* fr = frames + (size_t)b * 3 * L * kF;
    T A[kF];
#pragma unroll
    for (int e = 0; e < kF; ++e) A[e] = (e == 0 || e == 4 || e == 8) ? T(1) : T(0);
    bool bad = false;
    if (n > 0) {
        x[0] = x[1] = x[2] = T(0);
#pragma unroll
        for (int e = 0; e < kF; ++e) fr[e] = A[e];
    }
    for (int i = 1; i < 3 * n; ++i) {
        const int k = (i - 1) % 3, j = (i - 1) / 3;
        T Bm[kF], C[kF];
        plain_dihedral<T>(k, ang[(size_t)k * L + j], Bm);
        plain_compose<T>(A, Bm, C);
#pragma unroll
        for (int e = 0; e < kF; ++e) {
            A[e] = C[e];
            fr[(size_t)i * kF + e] = C[e];
        }
        x[3 * i] = C[9]; x[3 * i + 1] = C[10]; x[3 * i + 2] = C[11];
        bad |= (C[9] != C[9]) | (C[10] != C[10]) | (C[11] != C[11]);
    }
    for (size_t e = 9 * (size_t)n; e < 9 * (size_t)L; ++e) x[e] = T(0);
    if (bad) set_status(status, TPL_A2B_STATUS_NAN);
}

// One thread per (chain, angle row k, residue j): the reference per-pair
// formula dx_i/dtheta = A_{m-1} dB_m A_m^{-1} x_i (cBackboneProteinCUDAKernels.cu:84-111)
// with M = A_{m-1} dB_m A_m^{-1} hoisted, dotted with g_i and summed over
// i > m (cBackboneProteinCUDAKernels.cu:146-163).
template <typename T>
__global__ void a2b_bwd_plain(T* grad_angles, const T* grad_coords, const T* angles,
                              const int32_t* lengths, const T* frames, uint32_t* status,
                              int B, int L) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y, b = blockIdx.z;
    if (j >= L) return;
    int n = lengths[b];
    n = n < 0 ? 0 : (n > L ? L : n);
    T* out = grad_angles + ((size_t)b * 3 + k) * L + j;
    const int m = 3 * j + k + 1;
    if (j >= n || m >= 3 * n - 1) {
        *out = T(0);
        return;
    }
    const T* fr = frames + (size_t)b * 3 * L * kF;
    const T* g = grad_coords + (size_t)b * 9 * L;
    const T th = angles[((size_t)b * 3 + k) * L + j];
    const Bond bd = bond(k);
    T s, c;
    plain_sincos<T>(th, &s, &c);
    const T cb = (T)bd.cb, sb = (T)bd.sb;
    // dB rotation part by columns (Math/cMatrix44.cpp:54-59): column 0 is zero.
    const T D1[3] = {c * sb, -s, c * cb};
    const T D2[3] = {-s * sb, -c, -s * cb};
    const T* A1 = fr + (size_t)(m - 1) * kF;
    const T* A2 = fr + (size_t)m * kF;
    // K = A1.R * dB.R (columns), M_R = K * A2.R^T, M_t = -M_R * A2.t
    T K1[3], K2[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        K1[i] = A1[i] * D1[0] + A1[3 + i] * D1[1] + A1[6 + i] * D1[2];
        K2[i] = A1[i] * D2[0] + A1[3 + i] * D2[1] + A1[6 + i] * D2[2];
    }
    T MR[3][3];  // MR[i][c] = sum_q K_q[i] * A2.R[c][q]  (A2.R^T row q = column q)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) MR[i][cc] = K1[i] * A2[3 + cc] + K2[i] * A2[6 + cc];
    T Mt[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) Mt[i] = -(MR[i][0] * A2[9] + MR[i][1] * A2[10] + MR[i][2] * A2[11]);
    T sum = T(0);
    for (int i = m + 1; i < 3 * n; ++i) {
        const T* Ai = fr + (size_t)i * kF;
        const T x0 = Ai[9], x1 = Ai[10], x2 = Ai[11];
#pragma unroll
        for (int r = 0; r < 3; ++r)
            sum += g[3 * i + r] * (MR[r][0] * x0 + MR[r][1] * x1 + MR[r][2] * x2 + Mt[r]);
    }
    *out = sum;
    if (sum != sum) set_status(status, TPL_A2B_STATUS_NAN);
}


// ------------------------------------------------------------- host side
bool valid_config(const tpl_a2b_config* c) {
    if (!c) return false;
    const int f = c->fwd_res_per_lane, b = c->bwd_res_per_lane;
    const bool fok = (f == 1 || f == 2 || f == 4 || f == 8);
    const bool bok = (b == 1 || b == 2 || b == 4);
    return fok && bok && c->flags == 0 && c->reserved == 0;
}

// Residues per lane the registers hold without spilling: the forward keeps R
// angles per row and its frame, the backward also (cos, sin) of its 3R angles
// and its 9R coordinate gradients.
bool fits_dtype(int dtype, const tpl_a2b_config* c) {
    if (dtype == TPL_A2B_F32) return true;
    return c->fwd_res_per_lane <= 4 && c->bwd_res_per_lane <= 2;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
bool misaligned(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 255) != 0; }

int chunks(int L, int C) { return L <= 0 ? 0 : (L + C - 1) / C; }

// Forward -> backward state: one FP64 frame per backward chunk of every chain
// ([B][ncb][12] doubles; entry 0 of each chain, the identity, is not written).
size_t state_total(int B, int L, const tpl_a2b_config* c) {
    return align256((size_t)B * chunks(L, 64 * c->bwd_res_per_lane) * kF * sizeof(double));
}

// Sync: the forward's look-back words then the backward's (ticket_of,
// done_of and records_of give a chain's words).
struct SyncRegion {
    size_t tickets, done, pay, end;
    int sbx, prec;
};
// Tickets and done counters [8][sbx] (sbx = ceil(B / 8) rounded up to 64
// words), then records [B][prec] (prec = nc * pay rounded up to 32 words).
SyncRegion sync_region(size_t base, int B, int nc, int pay) {
    SyncRegion r{};
    const int bx = (B + 7) / 8;
    r.sbx = (bx + 63) / 64 * 64;
    r.prec = (int)(((int64_t)nc * pay + 31) / 32 * 32);
    r.tickets = base;
    r.done = r.tickets + (size_t)8 * r.sbx * 4;
    r.pay = align256(r.done + (size_t)8 * r.sbx * 4);
    r.end = align256(r.pay + (size_t)B * r.prec * sizeof(uint64_t));
    return r;
}
struct SyncLayout {
    SyncRegion fwd, bwd;
};
SyncLayout sync_layout(int B, int L, const tpl_a2b_config* c) {
    SyncLayout s{};
    s.fwd = sync_region(0, B, chunks(L, 64 * c->fwd_res_per_lane), kFwdPay);
    s.bwd = sync_region(s.fwd.end, B, chunks(L, 64 * c->bwd_res_per_lane), kBwdPay);
    return s;
}

using KernelFn = void (*)(ChunkArgs);

template <typename T>
KernelFn fwd_kernel(int r) {
    switch (r) {
        case 1: return a2b_fwd_chunk<T, 1>;
        case 2: return a2b_fwd_chunk<T, 2>;
        case 4: return a2b_fwd_chunk<T, 4>;
        default: return a2b_fwd_chunk<T, 8>;
    }
}
template <>
KernelFn fwd_kernel<double>(int r) {
    switch (r) {
        case 1: return a2b_fwd_chunk<double, 1>;
        case 2: return a2b_fwd_chunk<double, 2>;
        default: return a2b_fwd_chunk<double, 4>;
    }
}
// Chain kernels: W = slots per chain (1..kChainMax), dynamic LDS at most
// kChainMaxLds (W waves' staging), so that with the static LDS (sincos table,
// aggregates: under 1.5 KB) a workgroup stays within 64 KB.  The largest
// chain staging in use is 54 KB (FP32 forward with 8 residues per lane and 3
// slots, or FP64 with 4 and 3).
constexpr size_t kChainMaxLds = 56u << 10;
template <typename T, int R>
KernelFn fwd_chain_w(int w) {
    switch (w) {
        case 1: return a2b_fwd_chain<T, R, 1>;
        case 2: return a2b_fwd_chain<T, R, 2>;
        case 3: return a2b_fwd_chain<T, R, 3>;
        default: return a2b_fwd_chain<T, R, 4>;
    }
}
template <typename T>
KernelFn fwd_chain_kernel(int r, int w) {
    switch (r) {
        case 1: return fwd_chain_w<T, 1>(w);
        case 2: return fwd_chain_w<T, 2>(w);
        case 4: return fwd_chain_w<T, 4>(w);
        default: return fwd_chain_w<T, 8>(w);
    }
}
template <>
KernelFn fwd_chain_kernel<double>(int r, int w) {
    switch (r) {
        case 1: return fwd_chain_w<double, 1>(w);
        case 2: return fwd_chain_w<double, 2>(w);
        default: return fwd_chain_w<double, 4>(w);
    }
}
template <typename T, int R>
KernelFn bwd_chain_w(int w) {
    switch (w) {
        case 1: return a2b_bwd_chain<T, R, 1>;
        case 2: return a2b_bwd_chain<T, R, 2>;
        case 3: return a2b_bwd_chain<T, R, 3>;
        default: return a2b_bwd_chain<T, R, 4>;
    }
}
template <typename T>
KernelFn bwd_chain_kernel(int r, int w) {
    switch (r) {
        case 1: return bwd_chain_w<T, 1>(w);
        case 2: return bwd_chain_w<T, 2>(w);
        default: return bwd_chain_w<T, 4>(w);
    }
}
template <>
KernelFn bwd_chain_kernel<double>(int r, int w) {
    return r == 1 ? bwd_chain_w<double, 1>(w) : bwd_chain_w<double, 2>(w);
}
template <typename T>
KernelFn bwd_kernel(int r) {
    switch (r) {
        case 1: return a2b_bwd_chunk<T, 1>;
        case 2: return a2b_bwd_chunk<T, 2>;
        default: return a2b_bwd_chunk<T, 4>;
    }
}
template <>
KernelFn bwd_kernel<double>(int r) {
    return r == 1 ? a2b_bwd_chunk<double, 1> : a2b_bwd_chunk<double, 2>;
}

int check_common(int dtype, int B, int L) {
    if (dtype != TPL_A2B_F32 && dtype != TPL_A2B_F64) return TPL_A2B_ERR_ARG;
    if (B < 0 || L < 0) return TPL_A2B_ERR_ARG;
    // 9L and B*L index as int; the grid holds 8*ceil(B/8)*ceil(L/64) workgroups
    if ((int64_t)9 * L > INT32_MAX || (int64_t)B * L > INT32_MAX) return TPL_A2B_ERR_SIZE;
    if ((int64_t)8 * ((B + 7) / 8) * chunks(L, 64) > INT32_MAX) return TPL_A2B_ERR_SIZE;
    return TPL_A2B_OK;
}

ChunkArgs base_args(int B, int L, int C, const tpl_a2b_config* cfg) {
    ChunkArgs a{};
    a.B = B;
    a.L = L;
    a.nc = chunks(L, C);
    a.K = 64 * cfg->bwd_res_per_lane;
    a.nk = chunks(L, a.K);
    a.bx = (B + 7) / 8;
    return a;
}

void bind_sync(ChunkArgs& a, void* sync, const SyncRegion& r) {
    char* base = static_cast<char*>(sync);
    a.tickets = reinterpret_cast<uint32_t*>(base + r.tickets);
    a.done = reinterpret_cast<uint32_t*>(base + r.done);
    a.pay = reinterpret_cast<uint64_t*>(base + r.pay);
    a.sbx = r.sbx;
    a.prec = r.prec;
}

// The forward's look-back words for callers that cannot allocate
// (tpl_a2b_forward_noalloc): a file-scope device array, one copy per GPU,
// held until the process ends.  Host code must not pass `scratch` itself to
// a kernel (in host code the name is a stand-in, not a device address): it
// uses the address hipGetSymbolAddress returns on each device, taken once
// per device with that device current.
constexpr size_t kScratchBytes = 4u << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, or null on failure.
void* scratch_address() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

int launch(KernelFn k, const ChunkArgs& a, size_t lds, hipStream_t stream) {
    const dim3 grid((unsigned)(8 * a.bx * a.nc)), block(64);
    hipLaunchKernelGGL(k, grid, block, lds, stream, a);
    return hipGetLastError() == hipSuccess ? TPL_A2B_OK : TPL_A2B_ERR_LAUNCH;
}

// Workgroups of kernel k (threads each, dynamic LDS lds) that the device
// holds at once: the per-CU occupancy times the CUs, cached per device and
// kernel; 0 when the runtime cannot say.
int resident_workgroups(KernelFn k, int threads, size_t lds) {
    static std::mutex mu;
    static std::map<std::tuple<int, const void*, int, size_t>, int> cache;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    const auto key = std::make_tuple(dev, reinterpret_cast<const void*>(k), threads, lds);
    std::lock_guard<std::mutex> lock(mu);
    const auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k), threads, lds) !=
            hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    return cache[key] = per_cu * cus;
}

// The chain kernels take chains of at most kChainMax chunks whose waves'
// staging (lds each) fits, one workgroup of a.nc waves per chain, and only
// when every chain's workgroup is resident from the start.  Otherwise the
// chunk kernels, whose one-wave workgroups put the work first: at B=1024 of
// lengths 30-1000 with 4 residues per lane the backward's chain workgroups
// (218 VGPRs, two per CU) ran in two rounds behind waves with nothing to do,
// 16.3 -> 21.2 us; where they all fit at once they win (steps of fwd+bwd:
// 256 x 700 residues 24.6 -> 17.0 us, 512 x 50-500 22.7 -> 17.3 us, 32 x 200
// 14.9 -> 10.9 us, each at the best config of either form).
bool chain_fits(KernelFn k, const ChunkArgs& a, size_t lds) {
    if (a.nc > kChainMax || (size_t)a.nc * lds > kChainMaxLds) return false;
    return a.B <= resident_workgroups(k, 64 * a.nc, a.nc * lds);
}
int launch_chain(KernelFn k, const ChunkArgs& a, size_t lds, hipStream_t stream) {
    hipLaunchKernelGGL(k, dim3((unsigned)a.B), dim3(64 * a.nc), a.nc * lds, stream, a);
    return hipGetLastError() == hipSuccess ? TPL_A2B_OK : TPL_A2B_ERR_LAUNCH;
}

}  // namespace

// =================================================================== C ABI
extern "C" {

const char* tpl_a2b_strerror(int code) {
    switch (code) {
        case TPL_A2B_OK: return "ok";
        case TPL_A2B_ERR_ARG: return "invalid argument (null pointer, negative size, misaligned buffer or unknown dtype)";
        case TPL_A2B_ERR_CONFIG: return "unsupported chunk configuration";
        case TPL_A2B_ERR_LAUNCH: return "kernel launch failed";
        case TPL_A2B_ERR_HIP: return "HIP runtime call failed";
        case TPL_A2B_ERR_SIZE: return "batch * angles_stride too large";
        case TPL_A2B_ERR_SCRATCH:
            return "problem too large for the built-in scratch of tpl_a2b_forward_noalloc "
                   "(tpl_a2b_noalloc_max_bytes)";
        default: return "unknown error code";
    }
}

int tpl_a2b_default_config(int batch, int angles_stride, tpl_a2b_config* cfg) {
    if (!cfg || batch < 0 || angles_stride < 0) return TPL_A2B_ERR_ARG;
    cfg->fwd_res_per_lane = angles_stride <= 64 ? 1 : (angles_stride <= 128 ? 2 : 4);
    cfg->bwd_res_per_lane = angles_stride <= 64 ? 1 : 2;
    cfg->flags = 0;
    cfg->reserved = 0;
    return TPL_A2B_OK;
}

int64_t tpl_a2b_chunk_residues(const tpl_a2b_config* cfg, int backward) {
    if (!valid_config(cfg)) return -1;
    return (int64_t)64 * (backward ? cfg->bwd_res_per_lane : cfg->fwd_res_per_lane);
}

size_t tpl_a2b_state_bytes(int dtype, int batch, int angles_stride, const tpl_a2b_config* cfg) {
    if (check_common(dtype, batch, angles_stride) || !valid_config(cfg)) return 0;
    if (batch == 0 || angles_stride == 0) return 0;
    return state_total(batch, angles_stride, cfg);
}

size_t tpl_a2b_sync_bytes(int dtype, int batch, int angles_stride, const tpl_a2b_config* cfg) {
    if (check_common(dtype, batch, angles_stride) || !valid_config(cfg)) return 0;
    if (batch == 0 || angles_stride == 0) return 0;
    return sync_layout(batch, angles_stride, cfg).bwd.end;
}

int tpl_a2b_sync_clear(void* sync, size_t bytes, tpl_stream_t stream) {
    if (bytes == 0) return TPL_A2B_OK;
    if (!sync || misaligned(sync) || bytes % 4) return TPL_A2B_ERR_ARG;
    const size_t words = bytes / 4;
    const unsigned blocks = (unsigned)std::min<size_t>((words + 255) / 256, 1024);
    hipLaunchKernelGGL(a2b_zero_words, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<uint32_t*>(sync), words);
    return hipGetLastError() == hipSuccess ? TPL_A2B_OK : TPL_A2B_ERR_LAUNCH;
}

int tpl_a2b_forward(int dtype, const void* angles, void* coords, const int32_t* lengths,
                    void* state, void* sync, uint32_t* status, int batch, int angles_stride,
                    const tpl_a2b_config* cfg, tpl_stream_t stream) {
    int rc = check_common(dtype, batch, angles_stride);
    if (rc) return rc;
    if (!valid_config(cfg) || !fits_dtype(dtype, cfg)) return TPL_A2B_ERR_CONFIG;
    if (batch == 0 || angles_stride == 0) return TPL_A2B_OK;
    if (!angles || !coords || !lengths || !sync) return TPL_A2B_ERR_ARG;
    if (misaligned(state) || misaligned(sync)) return TPL_A2B_ERR_ARG;
    const int R = cfg->fwd_res_per_lane;
    ChunkArgs a = base_args(batch, angles_stride, 64 * R, cfg);
    a.angles = angles;
    a.coords = coords;
    a.lengths = lengths;
    a.status = status;
    a.ckpt = static_cast<double*>(state);
    bind_sync(a, sync, sync_layout(batch, angles_stride, cfg).fwd);
    const bool f32 = dtype == TPL_A2B_F32;
    const size_t lds = (size_t)9 * 64 * R * (f32 ? 4 : 8);
    if (a.nc <= kChainMax) {
        const KernelFn kc = f32 ? fwd_chain_kernel<float>(R, a.nc) : fwd_chain_kernel<double>(R, a.nc);
        if (chain_fits(kc, a, lds)) return launch_chain(kc, a, lds, static_cast<hipStream_t>(stream));
    }
    return launch(f32 ? fwd_kernel<float>(R) : fwd_kernel<double>(R), a, lds,
                  static_cast<hipStream_t>(stream));
}

int tpl_a2b_backward(int dtype, void* grad_angles, const void* grad_coords, const void* angles,
                     const int32_t* lengths, const void* state, void* sync, uint32_t* status,
                     int batch, int angles_stride, const tpl_a2b_config* cfg,
                     tpl_stream_t stream) {
    int rc = check_common(dtype, batch, angles_stride);
    if (rc) return rc;
    if (!valid_config(cfg) || !fits_dtype(dtype, cfg)) return TPL_A2B_ERR_CONFIG;
    if (batch == 0 || angles_stride == 0) return TPL_A2B_OK;
    if (!grad_angles || !grad_coords || !angles || !lengths || !state || !sync)
        return TPL_A2B_ERR_ARG;
    if (misaligned(state) || misaligned(sync)) return TPL_A2B_ERR_ARG;
    const int R = cfg->bwd_res_per_lane;
    ChunkArgs a = base_args(batch, angles_stride, 64 * R, cfg);
    a.angles = angles;
    a.grad = grad_coords;
    a.grad_angles = grad_angles;
    a.lengths = lengths;
    a.status = status;
    a.ckpt = const_cast<double*>(static_cast<const double*>(state));
    bind_sync(a, sync, sync_layout(batch, angles_stride, cfg).bwd);
    const bool f32 = dtype == TPL_A2B_F32;
    const size_t lds = (size_t)6 * R * 64 * sizeof(double);
    if (a.nc <= kChainMax) {
        const KernelFn kc = f32 ? bwd_chain_kernel<float>(R, a.nc) : bwd_chain_kernel<double>(R, a.nc);
        if (chain_fits(kc, a, lds)) return launch_chain(kc, a, lds, static_cast<hipStream_t>(stream));
    }
    return launch(f32 ? bwd_kernel<float>(R) : bwd_kernel<double>(R), a, lds,
                  static_cast<hipStream_t>(stream));
}

size_t tpl_a2b_noalloc_max_bytes(void) { return kScratchBytes; }

int tpl_a2b_forward_noalloc(int dtype, const void* angles, void* coords, const int32_t* lengths,
                            uint32_t* status, int batch, int angles_stride,
                            const tpl_a2b_config* cfg, tpl_stream_t stream) {
    int rc = check_common(dtype, batch, angles_stride);
    if (rc) return rc;
    if (!valid_config(cfg) || !fits_dtype(dtype, cfg)) return TPL_A2B_ERR_CONFIG;
    if (batch == 0 || angles_stride == 0) return TPL_A2B_OK;
    if (!angles || !coords || !lengths) return TPL_A2B_ERR_ARG;
    const SyncRegion r = sync_layout(batch, angles_stride, cfg).fwd;
    if (r.end > kScratchBytes) return TPL_A2B_ERR_SCRATCH;
    const int R = cfg->fwd_res_per_lane;
    ChunkArgs a = base_args(batch, angles_stride, 64 * R, cfg);
    a.angles = angles;
    a.coords = coords;
    a.lengths = lengths;
    a.status = status;
    a.ckpt = nullptr;
    const bool f32 = dtype == TPL_A2B_F32;
    const size_t lds = (size_t)9 * 64 * R * (f32 ? 4 : 8);
    // Chains of at most kChainMax chunks need no look-back words (the same
    // kernel as tpl_a2b_forward's).
    if (a.nc <= kChainMax) {
        const KernelFn kc = f32 ? fwd_chain_kernel<float>(R, a.nc) : fwd_chain_kernel<double>(R, a.nc);
        if (chain_fits(kc, a, lds)) return launch_chain(kc, a, lds, static_cast<hipStream_t>(stream));
    }
    void* base = scratch_address();
    if (!base) return TPL_A2B_ERR_HIP;
    // Tickets, done counters and records must start at zero.  Every call
    // leaves them zero unless it timed out, and a different shape lays them
    // out elsewhere, so the words this call uses are cleared first, with a
    // kernel (capture-safe).
    rc = tpl_a2b_sync_clear(base, r.end, stream);
    if (rc) return rc;
    bind_sync(a, base, r);
    return launch(f32 ? fwd_kernel<float>(R) : fwd_kernel<double>(R), a, lds,
                  static_cast<hipStream_t>(stream));
}

size_t tpl_a2b_plain_frames_bytes(int dtype, int batch, int angles_stride) {
    if (check_common(dtype, batch, angles_stride)) return 0;
    const size_t es = dtype == TPL_A2B_F32 ? 4 : 8;
    return (size_t)batch * 3 * angles_stride * kF * es;
}

int tpl_a2b_forward_plain(int dtype, const void* angles, void* coords, const int32_t* lengths,
                          void* frames, uint32_t* status, int batch, int angles_stride,
                          tpl_stream_t stream) {
    int rc = check_common(dtype, batch, angles_stride);
    if (rc) return rc;
    if (batch == 0 || angles_stride == 0) return TPL_A2B_OK;
    if (!angles || !coords || !lengths || !frames) return TPL_A2B_ERR_ARG;
    const dim3 grid((batch + 63) / 64), block(64);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (dtype == TPL_A2B_F32)
        hipLaunchKernelGGL(a2b_fwd_plain<float>, grid, block, 0, s,
                           static_cast<const float*>(angles), static_cast<float*>(coords),
                           lengths, static_cast<float*>(frames), status, batch, angles_stride);
    else
        hipLaunchKernelGGL(a2b_fwd_plain<double>, grid, block, 0, s,
                           static_cast<const double*>(angles), static_cast<double*>(coords),
                           lengths, static_cast<double*>(frames), status, batch, angles_stride);
    return hipGetLastError() == hipSuccess ? TPL_A2B_OK : TPL_A2B_ERR_LAUNCH;
}

int tpl_a2b_backward_plain(int dtype, void* grad_angles, const void* grad_coords,
                           const void* angles, const int32_t* lengths, const void* frames,
                           uint32_t* status, int batch, int angles_stride,
                           tpl_stream_t stream) {
    int rc = check_common(dtype, batch, angles_stride);
    if (rc) return rc;
    if (batch == 0 || angles_stride == 0) return TPL_A2B_OK;
    if (batch > 65535) return TPL_A2B_ERR_SIZE;
    if (!grad_angles || !grad_coords || !angles || !lengths || !frames) return TPL_A2B_ERR_ARG;
    const dim3 grid((angles_stride + 63) / 64, 3, batch), block(64);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (dtype == TPL_A2B_F32)
        hipLaunchKernelGGL(a2b_bwd_plain<float>, grid, block, 0, s,
                           static_cast<float*>(grad_angles), static_cast<const float*>(grad_coords),
                           static_cast<const float*>(angles), lengths,
                           static_cast<const float*>(frames), status, batch, angles_stride);
    else
        hipLaunchKernelGGL(a2b_bwd_plain<double>, grid, block, 0, s,
                           static_cast<double*>(grad_angles),
                           static_cast<const double*>(grad_coords),
                           static_cast<const double*>(angles), lengths,
                           static_cast<const double*>(frames), status, batch, angles_stride);
    return hipGetLastError() == hipSuccess ? TPL_A2B_OK : TPL_A2B_ERR_LAUNCH;
}

int tpl_a2b_tune(int dtype, const void* angles, const int32_t* lengths,
                 const void* grad_coords, void* coords, void* grad_angles, void* state,
                 void* sync, int batch, int angles_stride, const tpl_a2b_config* candidates,
                 int n_candidates, int reps, tpl_a2b_config* best, float* ms_per_candidate,
                 tpl_stream_t stream) {
    int rc = check_common(dtype, batch, angles_stride);
    if (rc) return rc;
    if (!candidates || n_candidates <= 0 || !best || reps <= 0) return TPL_A2B_ERR_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return TPL_A2B_ERR_HIP;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return TPL_A2B_ERR_HIP;
    }
    size_t sync_max = 0;  // the caller sized `sync` for the largest candidate
    std::vector<char> ok(n_candidates, 0);
    for (int c = 0; c < n_candidates; ++c) {
        ok[c] = valid_config(candidates + c) && fits_dtype(dtype, candidates + c);
        if (ok[c]) sync_max = std::max(sync_max, tpl_a2b_sync_bytes(dtype, batch, angles_stride, candidates + c));
    }
    // Steps are timed back to back, kStepsPerRun per event pair, as a caller's
    // loop runs them; the candidates take turns, pass by pass, so that a
    // drift of the clock or of the machine's load falls on all of them alike.
    // The first pass is the warm-up.  Every call leaves the look-back words
    // zero; clearing them before each run keeps one failed candidate from
    // affecting the next.
    constexpr int kStepsPerRun = 6;
    std::vector<std::vector<float>> times(n_candidates);
    rc = TPL_A2B_OK;
    for (int pass = 0; pass <= reps && rc == TPL_A2B_OK; ++pass) {
        for (int c = 0; c < n_candidates && rc == TPL_A2B_OK; ++c) {
            if (!ok[c]) continue;
            const tpl_a2b_config* cf = candidates + c;
            rc = tpl_a2b_sync_clear(sync, sync_max, stream);
            if (rc) break;
            if (hipEventRecord(e0, s) != hipSuccess) { rc = TPL_A2B_ERR_HIP; break; }
            for (int it = 0; it < kStepsPerRun && rc == TPL_A2B_OK; ++it) {
                rc = tpl_a2b_forward(dtype, angles, coords, lengths, state, sync, nullptr, batch,
                                     angles_stride, cf, stream);
                if (rc) break;
                rc = tpl_a2b_backward(dtype, grad_angles, grad_coords, angles, lengths, state, sync,
                                      nullptr, batch, angles_stride, cf, stream);
            }
            if (rc) break;
            if (hipEventRecord(e1, s) != hipSuccess) { rc = TPL_A2B_ERR_HIP; break; }
            if (hipEventSynchronize(e1) != hipSuccess) { rc = TPL_A2B_ERR_HIP; break; }
            float t = 0.f;
            if (hipEventElapsedTime(&t, e0, e1) != hipSuccess) { rc = TPL_A2B_ERR_HIP; break; }
            if (pass > 0) times[c].push_back(t / kStepsPerRun);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    float best_ms = -1.f;
    int best_i = -1;
    for (int c = 0; c < n_candidates; ++c) {
        float ms = -1.f;
        if (!times[c].empty()) {
            std::sort(times[c].begin(), times[c].end());
            ms = times[c][times[c].size() / 2];
            if (best_i < 0 || ms < best_ms) {
                best_ms = ms;
                best_i = c;
            }
        }
        if (ms_per_candidate) ms_per_candidate[c] = ms;
    }
    if (best_i < 0) return TPL_A2B_ERR_CONFIG;
    *best = candidates[best_i];
    return TPL_A2B_OK;
}

}  // extern "C"
