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
// Scan kernels (a2b_fwd_scan / a2b_bwd_scan):
//   * a tile is 64*W*R consecutive residues of one chain; lane l owns R of
//     them and moves its own angles, coordinates and gradients between HBM
//     and registers (whole runs in 16-byte accesses where aligned); lanes past
//     the chain write their zeros first;
//   * forward: lane l composes its R residues serially in FP64 (the structured
//     3x4 step below), the 64 lane aggregates are scanned with DPP (row_shr
//     1/2/4/8, row_bcast 15/31), the W wave aggregates through LDS, and tiles
//     of one chain through a decoupled look-back on flags in global memory
//     (tickets taken in chain order, so every tile waits only on tiles already
//     being processed); FP32 outputs also keep each coordinate's rounding
//     residual as a scaled half;
//   * backward: the O(L) adjoint from the forward's coordinates (rebuilt to
//     FP64 grade with the residuals),
//       dL/dtheta_j = u_j . (T_j - x_j x F_j),
//       F_j = sum_{i>j} g_i,  T_j = sum_{i>j} x_i x g_i,
//     with u_j the unit bond vector ending at atom j: a reverse (suffix) sum of
//     the 6-vectors (x_i x g_i, g_i) over lanes, waves and tiles.  No frame is
//     recomputed.
// Plain kernels (a2b_fwd_plain / a2b_bwd_plain) restate the reference GPU
// algorithm without the O(L^2) derivative tensor: they are the baseline.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
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


// ------------------------------------------------------------ cross-lane ops
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
// fences at agent scope:
//   * flags, tickets and payload words move with relaxed agent-scope atomic
//     loads/stores, which are coherent at device scope by themselves;
//   * a publisher orders its payload stores before the flag store with
//     s_waitcnt vmcnt(0) (vector memory operations complete), plus a
//     workgroup-scope fence that stops the compiler from reordering;
//   * a reader waits for its flag load the same way before loading payload.
constexpr int kWaitVmcnt0 = 0 | (0x7 << 4) | (0xf << 8);  // vmcnt(0); expcnt, lgkmcnt: no wait

__device__ __forceinline__ uint32_t flag_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void word_store(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Publishes flag value v after every earlier store of this thread completed.
__device__ __forceinline__ void flag_store(uint32_t* p, uint32_t v) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    word_store(p, v);
}
// After a flag was seen set: later loads are issued after the flag load.
__device__ __forceinline__ void after_flag() {
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
__device__ __forceinline__ void payload_store(double* p, double v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p),
                       static_cast<unsigned long long>(__double_as_longlong(v)), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double payload_load(const double* p) {
    return __longlong_as_double(static_cast<long long>(__hip_atomic_load(
        reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
}
__device__ __forceinline__ uint32_t atomic_inc(uint32_t* p) {
    return __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

constexpr uint32_t kFlagAggregate = 1, kFlagInclusive = 2;

// Polls a look-back flag until it is set.  The wait is bounded: after
// kSpinLimit polls (seconds, far beyond any legitimate wait) it reports
// TPL_A2B_STATUS_TIMEOUT and returns 0 so the caller skips that predecessor
// and the kernel still drains.
constexpr int kSpinLimit = 1 << 21;
__device__ __forceinline__ uint32_t wait_flag(const uint32_t* p, uint32_t* status) {
    for (int spins = 0; spins < kSpinLimit; ++spins) {
        const uint32_t f = flag_load(p);
        if (f != 0) {
            after_flag();
            return f;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    set_status(status, TPL_A2B_STATUS_TIMEOUT);
    return 0;
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

template <typename T, int N>
__device__ __forceinline__ void lane_store(T* p, int cnt, const T (&v)[N]) {
    constexpr int U = unit_bytes<T, N>();
    using V = typename Unit<U>::type;
    if (cnt >= N && (reinterpret_cast<uintptr_t>(p) % U) == 0) {
        V w[N * sizeof(T) / U];
        __builtin_memcpy(w, v, sizeof(v));
#pragma unroll
        for (int i = 0; i < (int)(N * sizeof(T) / U); ++i) reinterpret_cast<V*>(p)[i] = w[i];
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

// ------------------------------------------------------- FP32 output residual
// With FP32 coordinates the forward also keeps, per coordinate, the rounding
// residual of its FP64 value, r = x - (double)(float)x, as the half
// h = RTZ_f16(r * 2^24), in nine planes per chain: coordinate m (0..8) of
// residue r at res[(b*9 + m)*Lp + r], Lp = L rounded up to 8, so that the R
// halves of a lane in one plane are contiguous and consecutive lanes are
// adjacent.  The backward rebuilds x = (double)x32 + h * 2^-24,
// within |x| * 2^-36 of the FP64 value (|x| < 2^16: h stays finite), so the
// bond directions it takes from coordinate differences are FP64-grade; from
// the FP32 coordinates alone they would carry ~1e-5 relative error.
constexpr float kResScale = 0x1.0p24f, kResInv = 0x1.0p-24f;

// Scaled residual of x against its FP32 rounding f.
__device__ __forceinline__ float res_of(double x, float f) {
    return (float)(x - (double)f) * kResScale;
}
__device__ __forceinline__ uint32_t res_pack(float r0, float r1) {
    return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(r0, r1));
}
__device__ __forceinline__ double res_apply(float x32, _Float16 h) {
    return fma((double)((float)h), (double)kResInv, (double)x32);
}

// Look-back payload per (chain, tile): aggregate and inclusive value.  The
// forward uses 12 doubles each (frames), the backward 6 (sums).
constexpr int kFwdPayload = 2 * kF;
constexpr int kBwdPayload = 12;

struct ScanArgs {
    const void* angles;    // forward input
    const void* coords;    // backward input: the forward's output
    void* out;             // coords (forward) / grad_angles (backward)
    const void* grad;      // grad_coords (backward)
    const int32_t* lengths;
    uint32_t* status;
    _Float16* res;         // [B][9][res_lp] residuals of FP32 coordinates, or null
    int res_lp;            // halves per residual plane: L rounded up to 8
    int B, L;
    int ntf, trf;          // forward tiles per chain, residues per forward tile
    int ntb, trb;          // backward tiles per chain, residues per backward tile
    int bwhole;            // backward: one workgroup walks all tiles of a chain
    // forward look-back (ntf > 1), in the per-call scratch, zeroed before launch
    uint32_t* fflags;      // [B][ntf]
    uint32_t* ftickets;    // [B]
    double* fpay;          // [B][ntf][kFwdPayload]
    // backward look-back (ntb > 1), in the state: zeroed by the forward,
    // returned to zero by the backward
    uint32_t* bflags;      // [B][ntb]
    uint32_t* btickets;    // [B]
    uint32_t* bdone;       // [B]
    double* bpay;          // [B][ntb][kBwdPayload]
};

// The chain length as used: lengths outside [0, L] are clamped and reported.
__device__ __forceinline__ int clamp_length(const ScanArgs& a, int n, bool report) {
    if (n < 0 || n > a.L) {
        if (report) set_status(a.status, TPL_A2B_STATUS_LENGTH);
        n = n < 0 ? 0 : a.L;
    }
    return n;
}

// Work item -> tile of chain b = item % B.  With several tiles per chain, the
// tile comes from the chain's ticket counter, so a tile is taken only after
// every earlier ticket of its chain was taken by a running workgroup
// (forward: ascending tiles; backward: descending).
__device__ __forceinline__ void item_tile(int b, int nt, uint32_t* tickets, bool descending,
                                          int* s_tile, int* t_out) {
    int t = 0;
    if (nt > 1) {
        if (threadIdx.x == 0) *s_tile = (int)atomic_inc(tickets + b);
        __syncthreads();
        t = __builtin_amdgcn_readfirstlane(*s_tile);
        if (descending) t = nt - 1 - t;
    }
    *t_out = t;
}

// E <- tile_pfx * wave_aggr[0] * ... * wave_aggr[wave-1] * E, built right to
// left with in-place left products (E enters as the lane's exclusive prefix
// within its wave).
__device__ __forceinline__ void lane_prefix(const double (*s_tot)[kF], int wave,
                                            const double* tile_pfx, double E[kF]) {
    for (int w = wave - 1; w >= 0; --w) frame_lmul(s_tot[w], E);
    if (tile_pfx) frame_lmul(tile_pfx, E);
}

__device__ __forceinline__ void affine(const double E[kF], double p0, double p1, double p2,
                                       double x[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) x[i] = E[i] * p0 + E[3 + i] * p1 + E[6 + i] * p2 + E[9 + i];
}

__device__ __forceinline__ void cross(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
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

// ------------------------------------------------------------ forward scan
// One tile = TR = 64*W*R residues of chain b; lane l owns residues
// rl = r0 + l*R .. rl+R-1, i.e. chain atoms 3rl .. 3rl+3R-1 and output
// floats 9rl .. 9rl+9R-1.  Steps:
//   1. the lanes load their R angles per row straight into registers; waves
//      wholly past the chain skip to step 5;
//   2. pass 1: per lane, FP64 sincos and the serial product of its R residues;
//      local positions of the lane's atoms 1 .. 3R-1 go to the lane's LDS slot;
//   3. DPP scan over the wave, LDS over the waves, look-back over tiles;
//   4. pass 2: atom 0 is the translation of the lane's global prefix frame E,
//      atom a >= 1 is E p_a, written over p_a in LDS; FP32 outputs also write
//      their residual planes;
//   5. each wave stores its 9*64*R-float output span with 16-byte stores from
//      LDS, zeros past the chain.
// LB: chains span several tiles (look-back over tiles); AHEAD: one tile per
// chain on a grid smaller than the batch (load the next chain ahead).  The
// launch picks the variant, so each carries only the state it needs.
template <typename T, int R, int MAXT, bool LB, bool AHEAD>
__global__ void __launch_bounds__(MAXT) a2b_fwd_scan(ScanArgs a) {
    extern __shared__ __align__(16) unsigned char s_dyn[];
    constexpr int E9 = 9 * R;  // output floats per lane = the lane's LDS stride
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nw = blockDim.x >> 6;
    const int TR = blockDim.x * R;
    T* s_x = reinterpret_cast<T*>(s_dyn);
    T* s_l = s_x + tid * E9;              // this lane's slot
    const T* s_w = s_x + wave * 64 * E9;  // this wave's staging, in output order
    __shared__ double s_tab[128];
    __shared__ double s_tot[4][kF];
    __shared__ double s_pfx[kF];
    __shared__ double s_agg[kF];
    __shared__ int s_tile;
    const T* angles = static_cast<const T*>(a.angles);
    T* coords = static_cast<T*>(a.out);
    const int L = a.L;
    for (int i = tid; i < 128; i += blockDim.x) s_tab[i] = kTabCS[i];

    // With one tile per chain the items are independent, and a workgroup that
    // walks several of them (a grid smaller than B) loads the next chain's
    // length and angles while it works on the current one.  With several
    // tiles per chain the tile comes from the chain's ticket at the start of
    // the item, so nothing is loaded ahead.
    const int total = a.B * a.ntf;
    constexpr bool ahead = AHEAD && !LB;
    T ang[3][R];  // this lane's angles of the current item (rows phi, psi, omega)
    int n_ahead = 0;
    auto load_ahead = [&](int it, T (&dst)[3][R]) {
        const int nb = clamp_length(a, a.lengths[it % a.B], false);
        const int c = min(max(nb - tid * R, 0), R);
        if (wave * 64 * R < nb)
            lane_load_rows<T, R>(angles + (size_t)(it % a.B) * 3 * L + tid * R, L, c, L - tid * R, dst);
        return nb;
    };
    if (ahead && (int)blockIdx.x < total) n_ahead = load_ahead(blockIdx.x, ang);

    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        __syncthreads();  // s_tab written; LDS of the previous tile no longer read
        const int b = item % a.B;
        int t = 0, n;
        if constexpr (ahead) {
            n = n_ahead;
            if (tid == 0) clamp_length(a, a.lengths[b], true);  // reports a bad length
        } else {
            const int n_raw = a.lengths[b];  // issued before the ticket round trip
            if constexpr (LB) item_tile(b, a.ntf, a.ftickets, false, &s_tile, &t);
            n = clamp_length(a, n_raw, t == 0 && tid == 0);
        }
        const int r0 = t * TR;
        const int rl = r0 + tid * R;
        const int cnt = min(max(n - rl, 0), R);  // residues of the chain in this lane

        // Zero the backward's look-back words of the backward tiles inside
        // this forward tile (the backward returns them to zero after use).
        if (a.bflags) {
            const int per = TR / a.trb;
            for (int j = tid; j < per; j += blockDim.x) {
                const int tb = t * per + j;
                if (tb < a.ntb) word_store(a.bflags + (size_t)b * a.ntb + tb, 0u);
            }
            if (t == 0 && tid == 0) {
                word_store(a.btickets + b, 0u);
                word_store(a.bdone + b, 0u);
            }
        }

        const int rw0 = r0 + wave * 64 * R;  // first residue of the wave
        const bool wave_active = rw0 < n;    // uniform per wave
        const int next = item + (int)gridDim.x;
        T ang_next[3][R];
        double E[kF];
        if (wave_active) {
            if (!ahead) lane_load_rows<T, R>(angles + (size_t)b * 3 * L + rl, L, cnt, L - rl, ang);
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
        }
        // The next item's inputs load during this item's scan and stores.
        if (ahead && next < total) n_ahead = load_ahead(next, ang_next);
        if (wave_active) {
            wave_scan_frames(E);
            if (lane == 63) {
#pragma unroll
                for (int e = 0; e < kF; ++e) s_tot[wave][e] = E[e];
            }
            wave_shift_frames(E);
        } else if (lane == 63) {
#pragma unroll
            for (int e = 0; e < kF; ++e) s_tot[wave][e] = ident(e);
        }
        __syncthreads();
        if constexpr (LB) {
            if (tid == 0) {
                const size_t s0 = (size_t)b * a.ntf;
                uint32_t* fl = a.fflags + s0;
                double* pay = a.fpay + s0 * kFwdPayload;
                double* my = pay + (size_t)t * kFwdPayload;
                double acc[kF];
                {   // tile aggregate = product of the wave aggregates
#pragma unroll
                    for (int e = 0; e < kF; ++e) acc[e] = s_tot[nw - 1][e];
                    for (int w = nw - 2; w >= 0; --w) frame_lmul(s_tot[w], acc);
#pragma unroll
                    for (int e = 0; e < kF; ++e) {
                        s_agg[e] = acc[e];
                        payload_store(my + (t == 0 ? kF : 0) + e, acc[e]);
                    }
                }
                flag_store(fl + t, t == 0 ? kFlagInclusive : kFlagAggregate);
                frame_identity(acc);  // becomes this tile's exclusive prefix
                if (t > 0) {
                    for (int j = t - 1; j >= 0; --j) {
                        const uint32_t fj = wait_flag(fl + j, a.status);
                        if (fj == 0) continue;
                        const double* pj = pay + (size_t)j * kFwdPayload;
                        const bool incl = (fj == kFlagInclusive);
                        double v[kF];
#pragma unroll
                        for (int e = 0; e < kF; ++e) v[e] = payload_load(pj + (incl ? kF : 0) + e);
                        frame_lmul(v, acc);
                        if (incl) break;
                    }
                    double incl_v[kF];  // inclusive = prefix * aggregate
                    frame_rmul_to(acc, s_agg, incl_v);
#pragma unroll
                    for (int e = 0; e < kF; ++e) payload_store(my + kF + e, incl_v[e]);
                    flag_store(fl + t, kFlagInclusive);
                }
#pragma unroll
                for (int e = 0; e < kF; ++e) s_pfx[e] = acc[e];
            }
            __syncthreads();
        }
        T* x_b = coords + (size_t)b * 9 * L;
        const int g0 = 9 * rw0;  // the wave's output span [g0, g0 + 9*64*R) within [0, 9L)
        const int wcount = min(9 * 64 * R, 9 * L - g0);
        if (wave_active) {
            if (cnt > 0) {
                lane_prefix(s_tot, wave, LB ? s_pfx : nullptr, E);
                // Pass 2, in place in the lane's slot.  Residuals go to plane
                // m at residue rl + q as soon as a pair (q even, q + 1) is
                // complete (one half when R = 1): the lane's R halves per
                // plane are a whole run (rl + R <= Lp).
                [[maybe_unused]] float rpend[9];
                [[maybe_unused]] _Float16* rp = a.res + (size_t)b * 9 * a.res_lp + rl;
                const bool has_res = a.res != nullptr;  // uniform
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
                    const int q = at / 3;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const T o = (T)x[c];
                        s_l[3 * at + c] = o;
                        chk += (float)o;
                        if constexpr (sizeof(T) == 4) {
                            const int m = 3 * (at % 3) + c;  // plane
                            const float r = res_of(x[c], o);
                            if (has_res) {
                                _Float16* d = rp + (size_t)m * a.res_lp + q;
                                if constexpr (R == 1) {
                                    *reinterpret_cast<uint16_t*>(d) = (uint16_t)(res_pack(r, 0.f) & 0xffffu);
                                } else if (q % 2 == 0) {
                                    rpend[m] = r;
                                } else {
                                    *reinterpret_cast<uint32_t*>(d - 1) = res_pack(rpend[m], r);
                                }
                            }
                        }
                    }
                }
                if (!__builtin_isfinite(chk)) set_status(a.status, TPL_A2B_STATUS_NAN);
            }
            wave_lds_sync();
            wave_store_run<T>(x_b + g0, wcount, min(max(9 * n - g0, 0), wcount), s_w, lane);
        } else if (wcount > 0) {
            wave_store_run<T>(x_b + g0, wcount, 0, s_w, lane);
        }
        if (ahead) {
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int q = 0; q < R; ++q) ang[k][q] = ang_next[k][q];
        }
    }
}

// The lane's 9R coordinates as loaded: values (xr) and, for FP32, their
// residual halves (hr[9q + m] from plane m at residue rl + q; zero past the
// chain).
template <typename T, int R>
struct LaneCoords {
    T xr[9 * R];
    _Float16 hr[sizeof(T) == 4 ? 9 * R : 1];

    // x, g: the lane's first coordinate and coordinate gradient (into gr);
    // rp: plane 0 of the chain; rl: the lane's first residue; room: residues
    // of the row from rl.
    __device__ __forceinline__ void load(const T* x, const T* g, T (&gr)[9 * R], const _Float16* rp,
                                         int lp, int rl, int cnt, int room) {
        lane_load2<T, 9 * R>(x, g, 9 * cnt, 9 * room, xr, gr);
        if constexpr (sizeof(T) == 4) {
            // One 2R-byte load per plane, no branch: the run at residue rl is
            // 2R-aligned and inside the plane row for every lane of the row
            // (rl + R <= Lp); lanes past the row read the row's last run.
            // Halves past the chain are zeroed.
            const _Float16* p0 = rp + min(rl, lp - R);
#pragma unroll
            for (int m = 0; m < 9; ++m) {
                using V = typename Unit<2 * R>::type;
                const V u = *reinterpret_cast<const V*>(p0 + (size_t)m * lp);
                _Float16 h[R];
                __builtin_memcpy(h, &u, sizeof(h));
#pragma unroll
                for (int q = 0; q < R; ++q) hr[9 * q + m] = q < cnt ? h[q] : (_Float16)0;
            }
        }
    }
    // FP64 coordinates of the lane's atom `at` (0 .. 3R-1).
    __device__ __forceinline__ void atom(int at, double out[3]) const {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if constexpr (sizeof(T) == 4)
                out[c] = res_apply(xr[3 * at + c], hr[3 * at + c]);
            else
                out[c] = (double)xr[3 * at + c];
        }
    }
};

// ----------------------------------------------------------- backward scan
// The adjoint from the forward's coordinates.  Chain atom j >= 1 (atoms
// N_0, CA_0, C_0, N_1, ...) is placed by the transform whose angle theta_j is
// phi_r, psi_r, omega_r for j = 3r+1, 3r+2, 3r+3.  theta_j turns every later
// atom about the bond axis u_j = (x_j - x_{j-1}) / d_j through x_j, so
//   dL/dtheta_j = u_j . sum_{i>j} (x_i - x_j) x g_i = u_j . (T_j - x_j x F_j),
//   T_j = sum_{i>j} x_i x g_i,  F_j = sum_{i>j} g_i,
// a suffix sum of the 6-vectors (x_i x g_i, g_i).  Tiles of TR = 64*W*R
// residues are taken from the last tile of each chain to the first; lane l
// sums its own atoms 3rl .. 3rl+3R-1 and writes the gradients of its residues
// (atoms 3rl+1 .. 3rl+3R, the last one being the next lane's first).  Steps:
//   1. lanes wholly past the chain write their zero gradients at once; the
//      others load their coordinates and coordinate gradients (and, for FP32,
//      the residuals) into registers, zero past the chain, and keep them there
//      until pass B;
//   2. per lane, W = sum over its atoms of (x x g, g);
//   3. suffix sums of W over lanes (DPP), waves (LDS) and tiles (reverse
//      look-back) give C = the sum over all later atoms;
//   4. per atom, from the lane's last down: the suffix (T_j, F_j) and the
//      gradient; psi and omega of the last residue are 0.
template <typename T, int R, int MAXT>
__global__ void __launch_bounds__(MAXT) a2b_bwd_scan(ScanArgs a) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nw = blockDim.x >> 6;
    const int TR = blockDim.x * R;
    __shared__ double s_wsum[4][6];
    __shared__ double s_carry[6];
    __shared__ int s_tile;
    const T* xin = static_cast<const T*>(a.coords);
    const T* gin = static_cast<const T*>(a.grad);
    T* gang = static_cast<T*>(a.out);
    const int L = a.L;
    constexpr double kInvD[3] = {1.0 / 1.460, 1.0 / 1.525, 1.0 / 1.330};

    // Work items: with bwhole, a chain (its tiles walked from the last to the
    // first, the sum over later tiles carried in registers); else a tile,
    // chained to the later tiles of its chain by the reverse look-back.
    const bool whole = a.bwhole != 0;
    const int nitems = whole ? a.B : a.B * a.ntb;
    for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
        __syncthreads();
        const int b = item % a.B;
        const int n_raw = a.lengths[b];  // issued before the ticket round trip
        int t0 = 0;
        if (!whole) item_tile(b, a.ntb, a.btickets, true, &s_tile, &t0);
        const int n = clamp_length(a, n_raw, (whole || t0 == 0) && tid == 0);
        double carry[6] = {0, 0, 0, 0, 0, 0};  // whole: sum over the later tiles
        for (int step = 0; step < (whole ? a.ntb : 1); ++step) {
            const int t = whole ? a.ntb - 1 - step : t0;
            const int r0 = t * TR;
            const int rl = r0 + tid * R;
            const int cnt = min(max(n - rl, 0), R);
            const int crow = min(max(L - rl, 0), R);
            T* ga_l = gang + (size_t)b * 3 * L + rl;
            if (cnt == 0 && crow > 0) {
#pragma unroll
                for (int k = 0; k < 3; ++k) lane_store_zero<T, R>(ga_l + (size_t)k * L, crow);
            }
            // A walked tile wholly past the chain only writes its zeros
            // (uniform; it adds nothing to the carried sum).
            if (whole && r0 >= n) continue;
            if (whole) __syncthreads();  // s_wsum of the previous tile no longer read
            // Issue priority by the tiles still to walk after this one: the
            // longest chains are the kernel's critical path, and the waves of
            // shorter ones have slack (measured 21.3 -> 20.0 us at B=1024,
            // lengths 30-1000).
            switch (min(whole ? t : 0, 3)) {
                case 3: __builtin_amdgcn_s_setprio(3); break;
                case 2: __builtin_amdgcn_s_setprio(2); break;
                case 1: __builtin_amdgcn_s_setprio(1); break;
                default: __builtin_amdgcn_s_setprio(0); break;
            }
            const bool wave_active = r0 + wave * 64 * R < n;  // uniform per wave

            // Pass A: W = sum over the lane's atoms 0 .. 3R-1 (chain atoms
            // 3rl .. 3rl+3R-1) of (x x g, g); zero past the chain.
            const size_t o = (size_t)b * 9 * L + (size_t)9 * rl;
            const size_t ro = (size_t)b * 9 * a.res_lp + (size_t)rl;  // plane 0, residue rl
            double W[6] = {0, 0, 0, 0, 0, 0};
            LaneCoords<T, R> xc;
            T gr[9 * R];
            if (wave_active) {
                xc.load(xin + o, gin + o, gr, a.res ? a.res + (size_t)b * 9 * a.res_lp : nullptr,
                        a.res_lp, rl, cnt, L - rl);
#pragma unroll
                for (int at = 0; at < 3 * R; ++at) {
                    double x[3], xg[3];
                    xc.atom(at, x);
                    const double g[3] = {(double)gr[3 * at], (double)gr[3 * at + 1], (double)gr[3 * at + 2]};
                    cross(x, g, xg);
                    W[0] += xg[0]; W[1] += xg[1]; W[2] += xg[2];
                    W[3] += g[0]; W[4] += g[1]; W[5] += g[2];
                }
            }
            // C = sum over later lanes, waves and tiles.
            double C[6];
#pragma unroll
            for (int e = 0; e < 6; ++e) C[e] = W[e];
            if (wave_active) {
                wave_scan_sum<6>(C);  // inclusive prefix over lanes
#pragma unroll
                for (int e = 0; e < 6; ++e) {
                    const double tot = read_lane(C[e], 63);
                    if (lane == 0) s_wsum[wave][e] = tot;
                    C[e] = tot - C[e];
                }
            } else if (lane == 0) {
#pragma unroll
                for (int e = 0; e < 6; ++e) s_wsum[wave][e] = 0.0;
            }
            __syncthreads();
            for (int w = wave + 1; w < nw; ++w) {
#pragma unroll
                for (int e = 0; e < 6; ++e) C[e] += s_wsum[w][e];
            }
            if (whole) {
#pragma unroll
                for (int e = 0; e < 6; ++e) C[e] += carry[e];
                for (int w = 0; w < nw; ++w) {
#pragma unroll
                    for (int e = 0; e < 6; ++e) carry[e] += s_wsum[w][e];
                }
            } else if (a.ntb > 1) {  // then bflags, bpay, btickets, bdone are set (state)
                if (tid == 0) {
                    double tot[6] = {0, 0, 0, 0, 0, 0}, carry[6] = {0, 0, 0, 0, 0, 0};
                    for (int w = 0; w < nw; ++w) {
#pragma unroll
                        for (int e = 0; e < 6; ++e) tot[e] += s_wsum[w][e];
                    }
                    const int ntv = (n + TR - 1) / TR;  // tiles holding residues
                    if (t < ntv) {
                        const size_t s0 = (size_t)b * a.ntb;
                        uint32_t* fl = a.bflags + s0;
                        double* pay = a.bpay + s0 * kBwdPayload;
                        double* my = pay + (size_t)t * kBwdPayload;
                        if (t == ntv - 1) {
#pragma unroll
                            for (int e = 0; e < 6; ++e) payload_store(my + 6 + e, tot[e]);
                            flag_store(fl + t, kFlagInclusive);
                        } else {
#pragma unroll
                            for (int e = 0; e < 6; ++e) payload_store(my + e, tot[e]);
                            flag_store(fl + t, kFlagAggregate);
                            for (int j = t + 1; j < ntv; ++j) {
                                const uint32_t fj = wait_flag(fl + j, a.status);
                                if (fj == 0) continue;
                                const double* pj = pay + (size_t)j * kBwdPayload;
                                const bool incl = (fj == kFlagInclusive);
#pragma unroll
                                for (int e = 0; e < 6; ++e) carry[e] += payload_load(pj + (incl ? 6 : 0) + e);
                                if (incl) break;
                            }
#pragma unroll
                            for (int e = 0; e < 6; ++e) payload_store(my + 6 + e, carry[e] + tot[e]);
                            flag_store(fl + t, kFlagInclusive);
                        }
                    }
#pragma unroll
                    for (int e = 0; e < 6; ++e) s_carry[e] = carry[e];
                }
                __syncthreads();
#pragma unroll
                for (int e = 0; e < 6; ++e) C[e] += s_carry[e];
            }
            if (wave_active && cnt > 0) {
                // Pass B, on the lane's atoms still in registers from pass A,
                // plus chain atom 3(rl+R), the next lane's first, from memory.
                const bool has_next = rl + R < n;
                double xn[3], gn[3];  // chain atom 3(rl+R)
                // Without a next atom the loads read the lane's own first atom
                // (valid here, cnt > 0) and are discarded: no branch.
                const size_t on = has_next ? o + 9 * R : o;
                [[maybe_unused]] const size_t rn = has_next ? ro + R : ro;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const T xv = xin[on + c];
                    const T gv = gin[on + c];
                    double xd;
                    if constexpr (sizeof(T) == 4)
                        xd = res_apply((float)xv, a.res[rn + (size_t)c * a.res_lp]);
                    else
                        xd = (double)xv;
                    xn[c] = has_next ? xd : 0.0;
                    gn[c] = has_next ? (double)gv : 0.0;
                }
                // From the lane's atom 3R down: Xc = atom j, Xp = atom j-1, Xa =
                // atom j+1, S = sum over atoms after j of (x x g, g).  Atom 3R
                // belongs to the next lane, whose sum is inside C.
                double S[6], Xc[3], Xp[3], Xa[3];
                {
                    double xg[3];
                    cross(xn, gn, xg);
                    S[0] = C[0] - xg[0]; S[1] = C[1] - xg[1]; S[2] = C[2] - xg[2];
                    S[3] = C[3] - gn[0]; S[4] = C[4] - gn[1]; S[5] = C[5] - gn[2];
                }
#pragma unroll
                for (int c = 0; c < 3; ++c) Xc[c] = xn[c];
                xc.atom(3 * R - 1, Xp);
                T out[3][R];
#pragma unroll
                for (int j = 3 * R; j >= 1; --j) {
                    if (j < 3 * R) {
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            Xa[c] = Xc[c];
                            Xc[c] = Xp[c];
                        }
                        xc.atom(j - 1, Xp);
                    }
                    if (j == 3 * R - 1) {
#pragma unroll
                        for (int e = 0; e < 6; ++e) S[e] = C[e];
                    } else if (j < 3 * R - 1) {
                        const double g[3] = {(double)gr[3 * j + 3], (double)gr[3 * j + 4], (double)gr[3 * j + 5]};
                        double xg[3];
                        cross(Xa, g, xg);
                        S[0] += xg[0]; S[1] += xg[1]; S[2] += xg[2];
                        S[3] += g[0]; S[4] += g[1]; S[5] += g[2];
                    }
                    const int k = (j - 1) % 3, q = (j - 1) / 3;
                    double u[3], xf[3];
#pragma unroll
                    for (int c = 0; c < 3; ++c) u[c] = (Xc[c] - Xp[c]) * kInvD[k];
                    cross(Xc, S + 3, xf);
                    const double v = u[0] * (S[0] - xf[0]) + u[1] * (S[1] - xf[1]) + u[2] * (S[2] - xf[2]);
                    const int r = rl + q;
                    const bool zero = (r >= n) || (r == n - 1 && k > 0);
                    out[k][q] = zero ? (T)0 : (T)v;
                }
                bool bad = false;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
#pragma unroll
                    for (int q = 0; q < R; ++q) bad |= (out[k][q] != out[k][q]);
                    lane_store<T, R>(ga_l + (size_t)k * L, crow, out[k]);
                }
                if (bad) set_status(a.status, TPL_A2B_STATUS_NAN);
            }

        }  // tiles of a whole chain
        // Last tile of the chain to finish returns its look-back words to 0.
        if (!whole && a.ntb > 1 && tid == 0) {
            if (atomic_inc(a.bdone + b) == (uint32_t)(a.ntb - 1)) {
                for (int j = 0; j < a.ntb; ++j) word_store(a.bflags + (size_t)b * a.ntb + j, 0u);
                word_store(a.btickets + b, 0u);
                word_store(a.bdone + b, 0u);
            }
        }
    }
}

__global__ void a2b_zero_words(uint32_t* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
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
    const int w = c->waves, r = c->res_per_lane;
    const bool wok = (w == 1 || w == 2 || w == 4);
    const bool rok = (r == 1 || r == 2 || r == 4 || r == 8);
    return wok && rok && c->max_wgs >= 0 && c->flags == 0;
}

// Largest tiles (each lane's runs stay in registers, no spills), and the
// backward's residues per lane: one, so that a lane keeps its coordinates,
// gradients and residuals in registers from pass A to pass B (about 100
// VGPRs, four or more waves per SIMD).
constexpr int kMaxTileF32 = 1024, kMaxTileF64 = 512;
constexpr int kBwdR = 1;

bool fits_dtype(int dtype, const tpl_a2b_config* c) {
    const int tr = 64 * c->waves * c->res_per_lane;
    return tr <= (dtype == TPL_A2B_F32 ? kMaxTileF32 : kMaxTileF64);
}

// Tiling of one call: forward tiles of trf = 64*W*R residues; backward tiles
// of trb = 64*W*kBwdR residues, a divisor of trf.
struct Tiling {
    int trf, ntf, trb, ntb, rb;
};

Tiling tiling(int dtype, int L, const tpl_a2b_config* c) {
    Tiling t{};
    t.trf = 64 * c->waves * c->res_per_lane;
    t.rb = kBwdR;
    t.trb = 64 * c->waves * t.rb;
    t.ntf = L <= 0 ? 0 : (L + t.trf - 1) / t.trf;
    t.ntb = L <= 0 ? 0 : (L + t.trb - 1) / t.trb;
    return t;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// state and sync regions start on 256-byte boundaries (the kernels move the
// residual runs in aligned 4- to 16-byte units).
bool misaligned(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 255) != 0; }

// Per-call scratch of the forward's own look-back (ntf > 1): flags [B][ntf]
// and tickets [B] (zeroed by a2b_zero_words before the forward), payload.
struct SyncLayout {
    size_t flags_off, tickets_off, words, pay_off, total;
};
SyncLayout sync_layout(int B, int ntf) {
    SyncLayout s{};
    if (ntf <= 1) return s;
    const size_t tiles = (size_t)B * ntf;
    s.flags_off = 0;
    s.tickets_off = tiles * 4;
    s.words = tiles + B;
    s.pay_off = align256(s.words * 4);
    s.total = s.pay_off + tiles * kFwdPayload * sizeof(double);
    return s;
}

// Forward -> backward state: FP32 coordinate residuals [B][9][Lp] halves,
// Lp = L rounded up to 8 (FP32 only), then the backward's look-back words
// and payload (ntb > 1).
int res_lp(int L) { return (L + 7) & ~7; }

struct StateLayout {
    size_t res_off, pay_off, flags_off, tickets_off, done_off, total;
};
StateLayout state_layout(int dtype, int B, int L, int ntb) {
    StateLayout s{};
    s.res_off = 0;
    const size_t res = dtype == TPL_A2B_F32 ? (size_t)B * 9 * res_lp(L) * 2 : 0;
    s.total = align256(res);
    if (ntb <= 1) return s;
    const size_t tiles = (size_t)B * ntb;
    s.pay_off = s.total;
    s.flags_off = s.pay_off + tiles * kBwdPayload * sizeof(double);
    s.tickets_off = s.flags_off + tiles * 4;
    s.done_off = s.tickets_off + (size_t)B * 4;
    s.total = align256(s.done_off + (size_t)B * 4);
    return s;
}

using KernelFn = void (*)(ScanArgs);

constexpr int kMaxThreads = 256;  // up to 4 waves per workgroup
constexpr int kWholeMaxTiles = 4, kWholeMinChains = 512;

// Forward variants: (LB, AHEAD) = (1, 0) several tiles per chain, (0, 0) one
// tile per chain and one workgroup each, (0, 1) one tile per chain on a grid
// smaller than the batch.
template <typename T, int R>
KernelFn pick_kernel(bool fwd, bool lb, bool ahead) {
    if (!fwd) return a2b_bwd_scan<T, kBwdR, kMaxThreads>;
    if (lb) return a2b_fwd_scan<T, R, kMaxThreads, true, false>;
    return ahead ? a2b_fwd_scan<T, R, kMaxThreads, false, true>
                 : a2b_fwd_scan<T, R, kMaxThreads, false, false>;
}

template <typename T>
KernelFn pick_kernel(bool fwd, int r, bool lb, bool ahead) {
    switch (r) {
        case 1: return pick_kernel<T, 1>(fwd, lb, ahead);
        case 2: return pick_kernel<T, 2>(fwd, lb, ahead);
        case 4: return pick_kernel<T, 4>(fwd, lb, ahead);
        default: return pick_kernel<T, 8>(fwd, lb, ahead);
    }
}

int check_common(int dtype, int B, int L) {
    if (dtype != TPL_A2B_F32 && dtype != TPL_A2B_F64) return TPL_A2B_ERR_ARG;
    if (B < 0 || L < 0) return TPL_A2B_ERR_ARG;
    if ((int64_t)9 * L > INT32_MAX || (int64_t)B * L > INT32_MAX) return TPL_A2B_ERR_SIZE;
    return TPL_A2B_OK;
}

// state == nullptr (forward only): no residuals, no backward words.
int launch_scan(bool fwd, int dtype, ScanArgs args, void* state, void* sync,
                const tpl_a2b_config* cfg, hipStream_t stream) {
    const Tiling tl = tiling(dtype, args.L, cfg);
    args.trf = tl.trf;
    args.ntf = tl.ntf;
    args.trb = tl.trb;
    args.ntb = tl.ntb;
    args.res_lp = res_lp(args.L);
    if (state) {
        const StateLayout st = state_layout(dtype, args.B, args.L, tl.ntb);
        char* base = static_cast<char*>(state);
        if (dtype == TPL_A2B_F32) args.res = reinterpret_cast<_Float16*>(base + st.res_off);
        if (tl.ntb > 1) {
            args.bpay = reinterpret_cast<double*>(base + st.pay_off);
            args.bflags = reinterpret_cast<uint32_t*>(base + st.flags_off);
            args.btickets = reinterpret_cast<uint32_t*>(base + st.tickets_off);
            args.bdone = reinterpret_cast<uint32_t*>(base + st.done_off);
        }
    }
    if (fwd && tl.ntf > 1) {
        const SyncLayout sl = sync_layout(args.B, tl.ntf);
        char* base = static_cast<char*>(sync);
        args.fflags = reinterpret_cast<uint32_t*>(base + sl.flags_off);
        args.ftickets = reinterpret_cast<uint32_t*>(base + sl.tickets_off);
        args.fpay = reinterpret_cast<double*>(base + sl.pay_off);
        const int zb = (int)std::min<size_t>((sl.words + 255) / 256, 1024);
        hipLaunchKernelGGL(a2b_zero_words, dim3(zb), dim3(256), 0, stream, args.fflags,
                           (int)sl.words);
        if (hipGetLastError() != hipSuccess) return TPL_A2B_ERR_LAUNCH;
    }
    // Backward: with few tiles per chain and enough chains to fill the device,
    // one workgroup takes a whole chain (no tickets, no look-back waits);
    // otherwise one workgroup per tile.
    args.bwhole = (!fwd && tl.ntb > 1 && tl.ntb <= kWholeMaxTiles && args.B >= kWholeMinChains) ? 1 : 0;
    const int nt = fwd ? tl.ntf : (args.bwhole ? 1 : tl.ntb);
    int grid = args.B * nt;
    if (fwd && cfg->max_wgs > 0) grid = std::min(grid, cfg->max_wgs);
    const int threads = 64 * cfg->waves;
    const int r = fwd ? cfg->res_per_lane : tl.rb;
    const bool f32 = dtype == TPL_A2B_F32;
    const bool lb = fwd && tl.ntf > 1;
    const bool ahead = fwd && !lb && grid < args.B;
    KernelFn k = f32 ? pick_kernel<float>(fwd, r, lb, ahead) : pick_kernel<double>(fwd, r, lb, ahead);
    // The forward stages each lane's 9R coordinates in LDS, in output order.
    const size_t lds = fwd ? (size_t)(9 * r) * (f32 ? 4 : 8) * threads : 0;
    hipLaunchKernelGGL(k, dim3(grid), dim3(threads), lds, stream, args);
    return hipGetLastError() == hipSuccess ? TPL_A2B_OK : TPL_A2B_ERR_LAUNCH;
}

}  // namespace

// =================================================================== C ABI
extern "C" {

const char* tpl_a2b_strerror(int code) {
    switch (code) {
        case TPL_A2B_OK: return "ok";
        case TPL_A2B_ERR_ARG: return "invalid argument (null pointer, negative size or unknown dtype)";
        case TPL_A2B_ERR_CONFIG: return "unsupported tile configuration";
        case TPL_A2B_ERR_LAUNCH: return "kernel launch failed";
        case TPL_A2B_ERR_HIP: return "HIP runtime call failed";
        case TPL_A2B_ERR_SIZE: return "batch * angles_stride too large";
        default: return "unknown error code";
    }
}

int tpl_a2b_default_config(int batch, int angles_stride, tpl_a2b_config* cfg) {
    if (!cfg || batch < 0 || angles_stride < 0) return TPL_A2B_ERR_ARG;
    cfg->waves = 4;
    cfg->max_wgs = 0;
    cfg->flags = 0;
    // One tile per chain up to 1024 residues, else 1024-residue tiles.
    const int L = angles_stride;
    cfg->res_per_lane = L <= 256 ? 1 : (L <= 512 ? 2 : 4);
    return TPL_A2B_OK;
}

int64_t tpl_a2b_tile_residues(const tpl_a2b_config* cfg) {
    if (!valid_config(cfg)) return -1;
    return (int64_t)64 * cfg->waves * cfg->res_per_lane;
}

size_t tpl_a2b_state_bytes(int dtype, int batch, int angles_stride, const tpl_a2b_config* cfg) {
    if (check_common(dtype, batch, angles_stride) || !valid_config(cfg)) return 0;
    if (batch == 0 || angles_stride == 0) return 0;
    return state_layout(dtype, batch, angles_stride, tiling(dtype, angles_stride, cfg).ntb).total;
}

size_t tpl_a2b_sync_bytes(int dtype, int batch, int angles_stride, const tpl_a2b_config* cfg) {
    if (check_common(dtype, batch, angles_stride) || !valid_config(cfg)) return 0;
    if (batch == 0 || angles_stride == 0) return 0;
    return sync_layout(batch, tiling(dtype, angles_stride, cfg).ntf).total;
}

int tpl_a2b_forward(int dtype, const void* angles, void* coords, const int32_t* lengths,
                    void* state, void* sync, uint32_t* status, int batch, int angles_stride,
                    const tpl_a2b_config* cfg, tpl_stream_t stream) {
    int rc = check_common(dtype, batch, angles_stride);
    if (rc) return rc;
    if (!valid_config(cfg) || !fits_dtype(dtype, cfg)) return TPL_A2B_ERR_CONFIG;
    if (batch == 0 || angles_stride == 0) return TPL_A2B_OK;
    if (!angles || !coords || !lengths) return TPL_A2B_ERR_ARG;
    if (tpl_a2b_sync_bytes(dtype, batch, angles_stride, cfg) && !sync) return TPL_A2B_ERR_ARG;
    if (misaligned(state) || misaligned(sync)) return TPL_A2B_ERR_ARG;
    ScanArgs a{};
    a.angles = angles;
    a.out = coords;
    a.lengths = lengths;
    a.status = status;
    a.B = batch;
    a.L = angles_stride;
    void* st = tpl_a2b_state_bytes(dtype, batch, angles_stride, cfg) ? state : nullptr;
    return launch_scan(true, dtype, a, st, sync, cfg, static_cast<hipStream_t>(stream));
}

int tpl_a2b_backward(int dtype, void* grad_angles, const void* grad_coords, const void* coords,
                     const int32_t* lengths, void* state, uint32_t* status, int batch,
                     int angles_stride, const tpl_a2b_config* cfg, tpl_stream_t stream) {
    int rc = check_common(dtype, batch, angles_stride);
    if (rc) return rc;
    if (!valid_config(cfg) || !fits_dtype(dtype, cfg)) return TPL_A2B_ERR_CONFIG;
    if (batch == 0 || angles_stride == 0) return TPL_A2B_OK;
    if (!grad_angles || !grad_coords || !coords || !lengths) return TPL_A2B_ERR_ARG;
    const bool need_state = tpl_a2b_state_bytes(dtype, batch, angles_stride, cfg) != 0;
    if (need_state && !state) return TPL_A2B_ERR_ARG;
    if (need_state && misaligned(state)) return TPL_A2B_ERR_ARG;
    ScanArgs a{};
    a.coords = coords;
    a.out = grad_angles;
    a.grad = grad_coords;
    a.lengths = lengths;
    a.status = status;
    a.B = batch;
    a.L = angles_stride;
    return launch_scan(false, dtype, a, need_state ? state : nullptr, nullptr, cfg,
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
    float best_ms = -1.f;
    int best_i = -1;
    rc = TPL_A2B_OK;
    std::vector<float> times;
    for (int c = 0; c < n_candidates && rc == TPL_A2B_OK; ++c) {
        const tpl_a2b_config* cf = candidates + c;
        float ms = -1.f;
        if (valid_config(cf) && fits_dtype(dtype, cf)) {
            times.clear();
            for (int it = 0; it <= reps && rc == TPL_A2B_OK; ++it) {
                if (hipEventRecord(e0, s) != hipSuccess) { rc = TPL_A2B_ERR_HIP; break; }
                rc = tpl_a2b_forward(dtype, angles, coords, lengths, state, sync, nullptr, batch,
                                     angles_stride, cf, stream);
                if (rc) break;
                rc = tpl_a2b_backward(dtype, grad_angles, grad_coords, coords, lengths, state,
                                      nullptr, batch, angles_stride, cf, stream);
                if (rc) break;
                if (hipEventRecord(e1, s) != hipSuccess) { rc = TPL_A2B_ERR_HIP; break; }
                if (hipEventSynchronize(e1) != hipSuccess) { rc = TPL_A2B_ERR_HIP; break; }
                float t = 0.f;
                if (hipEventElapsedTime(&t, e0, e1) != hipSuccess) { rc = TPL_A2B_ERR_HIP; break; }
                if (it > 0) times.push_back(t);  // the first run is the warm-up
            }
            if (rc == TPL_A2B_OK && !times.empty()) {
                std::sort(times.begin(), times.end());
                ms = times[times.size() / 2];
                if (best_i < 0 || ms < best_ms) {
                    best_ms = ms;
                    best_i = c;
                }
            }
        }
        if (ms_per_candidate) ms_per_candidate[c] = ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (best_i < 0) return TPL_A2B_ERR_CONFIG;
    *best = candidates[best_i];
    return TPL_A2B_OK;
}

}  // extern "C"
