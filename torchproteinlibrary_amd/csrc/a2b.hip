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
//   * a tile is 64*W*R consecutive residues of one chain; its angles, coordinates
//     and coordinate gradients pass through LDS so that every global access is
//     a coalesced row segment;
//   * lane l composes its R residues serially in FP64 (the structured 3x4 step
//     below), the 64 lane aggregates are scanned with DPP (row_shr 1/2/4/8,
//     row_bcast 15/31), the W wave aggregates through LDS, and tiles of one
//     chain through a decoupled look-back on flags in global memory (tickets
//     taken in chain order, so every tile waits only on tiles already being
//     processed);
//   * the backward recomputes frames from the forward's per-tile FP64 prefix
//     checkpoints and runs the O(L) adjoint
//       dL/dtheta_m = u_m . T_m + (x_m x u_m) . F_m,
//       F_m = sum_{i>m} g_i,  T_m = sum_{i>m} x_i x g_i,
//     a reverse (suffix) sum of 6-vector wrenches (x_i x g_i, g_i) over lanes,
//     waves and tiles; u_m is column 0 of frame m (the bond axis rotated by
//     theta_m).
// Plain kernels (a2b_fwd_plain / a2b_bwd_plain) restate the reference GPU
// algorithm without the O(L^2) derivative tensor: they are the baseline.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

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

// f <- M * f, in place one column at a time (M applied after f).  M may live
// in registers, LDS or global memory.
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

// ------------------------------------------------------------------ sincos
// sin and cos of x in FP64 without library calls on the common path.
// Cody-Waite reduction x = k*pi/2 + r with pi/2 split into three parts: two
// parts of 33 significant bits, so k*P1 and k*P2 are exact for |k| < 2^20,
// plus the rounded remainder.  Then Taylor series on |r| <= pi/4: sin to r^15
// and cos to r^16.  The truncation error is below 5e-17, so the result is within
// ~2 ulp of the true value for |x| < 1.6e6.  Larger |x| fall back to the
// library's sincos.
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
constexpr double kPio2_1 = 0x1.921fb54400000p+0;
constexpr double kPio2_2 = 0x1.0b4611a600000p-34;
constexpr double kPio2_3 = 0x1.3198a2e037073p-69;
// (-1)^i / (2i+1)!, i = 1..7 and (-1)^i / (2i)!, i = 1..8, rounded to double.
constexpr double kSin[7] = {-0x1.5555555555555p-3, 0x1.1111111111111p-7, -0x1.a01a01a01a01ap-13,
                            0x1.71de3a556c734p-19, -0x1.ae64567f544e4p-26, 0x1.6124613a86d09p-33,
                            -0x1.ae7f3e733b81fp-41};
constexpr double kCos[8] = {-0x1.0000000000000p-1, 0x1.5555555555555p-5, -0x1.6c16c16c16c17p-10,
                            0x1.a01a01a01a01ap-16, -0x1.27e4fb7789f5cp-22, 0x1.1eed8eff8d898p-29,
                            -0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-45};

__device__ __forceinline__ void sincos_fp64(double x, double* s_out, double* c_out) {
    const double k = rint(x * kTwoOverPi);
    if (!(fabs(k) < 1048576.0)) {  // also catches NaN/inf
        sincos(x, s_out, c_out);
        return;
    }
    double r = fma(-k, kPio2_1, x);
    r = fma(-k, kPio2_2, r);
    r = fma(-k, kPio2_3, r);
    const double r2 = r * r;
    double ps = kSin[6];
#pragma unroll
    for (int i = 5; i >= 0; --i) ps = fma(ps, r2, kSin[i]);
    double pc = kCos[7];
#pragma unroll
    for (int i = 6; i >= 0; --i) pc = fma(pc, r2, kCos[i]);
    const double sr = fma(r * r2, ps, r);
    const double cr = fma(r2, pc, 1.0);
    const int q = static_cast<int>(k) & 3;
    const double s = (q & 1) ? cr : sr;
    const double c = (q & 1) ? sr : cr;
    *s_out = (q & 2) ? -s : s;
    *c_out = ((q + 1) & 2) ? -c : c;
}

// f <- f * B(angle) for row K of the tile's staged angles.
template <int K, typename T>
__device__ __forceinline__ void step_angle(double f[kF], T angle) {
    double s, c;
    sincos_fp64(static_cast<double>(angle), &s, &c);
    frame_step<K>(f, c, s);
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

// One Hillis-Steele step: f <- (frame of the source lane) * f; lanes that
// receive nothing compose with the identity.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void frame_scan_step(double f[kF]) {
    double g[kF];
#pragma unroll
    for (int e = 0; e < kF; ++e) g[e] = dpp_move<CTRL, ROW_MASK>(f[e], ident(e));
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

// Inclusive prefix sum over the wave of N doubles.
template <int N>
__device__ __forceinline__ void wave_scan_sum(double v[N]) {
#pragma unroll
    for (int e = 0; e < N; ++e) {
        double x = v[e];
        x += dpp_move<kRowShr1, 0xf>(x, 0.0);
        x += dpp_move<kRowShr2, 0xf>(x, 0.0);
        x += dpp_move<kRowShr4, 0xf>(x, 0.0);
        x += dpp_move<kRowShr8, 0xf>(x, 0.0);
        x += dpp_move<kRowBcast15, 0xa>(x, 0.0);
        x += dpp_move<kRowBcast31, 0xc>(x, 0.0);
        v[e] = x;
    }
}

// --------------------------------------------------------------- I/O helpers
__device__ __forceinline__ void set_status(uint32_t* status, uint32_t bits) {
    if (status) __hip_atomic_fetch_or(status, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t flag_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void flag_store(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
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
        if (f != 0) return f;
        __builtin_amdgcn_s_sleep(1);
    }
    set_status(status, TPL_A2B_STATUS_TIMEOUT);
    return 0;
}

// Look-back payload per (chain, tile): aggregate and inclusive value, 12
// doubles each (frames in the forward, wrenches use the first 6).
constexpr int kPayload = 2 * kF;

struct ScanArgs {
    const void* angles;
    void* out;             // coords (forward) / grad_angles (backward)
    const void* grad;      // grad_coords (backward)
    const int32_t* lengths;
    double* state;         // [B][ntiles][12] prefix checkpoints (ntiles > 1)
    uint32_t* flags;       // [B][ntiles] look-back flags, zero at launch
    uint32_t* ticket;      // tile ticket counter, zero at launch
    double* payload;       // [B][ntiles][kPayload]
    uint32_t* status;
    int B, L, ntiles;
    int lookback;          // chains may span several tiles
};

__device__ __forceinline__ int clamp_length(const ScanArgs& a, int b, bool report) {
    int n = a.lengths[b];
    if (n < 0 || n > a.L) {
        if (report) set_status(a.status, TPL_A2B_STATUS_LENGTH);
        n = n < 0 ? 0 : a.L;
    }
    return n;
}

// Next tile of this workgroup: a ticket (look-back mode, chain order) or a
// grid-stride index.  Returns -1 when no work is left.
__device__ __forceinline__ int next_tile(const ScanArgs& a, int iter, int* s_tile) {
    const int total = a.B * a.ntiles;
    __syncthreads();  // LDS of the previous tile is no longer read
    int tile;
    if (a.lookback) {
        if (threadIdx.x == 0)
            *s_tile = (int)__hip_atomic_fetch_add(a.ticket, 1u, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        tile = __builtin_amdgcn_readfirstlane(*s_tile);
    } else {
        tile = blockIdx.x + iter * gridDim.x;
    }
    return tile < total ? tile : -1;
}

// Dynamic LDS: the tile's angles [3][TR] and atom xyz [9*TR] (coords or
// grad_coords) in the I/O type, then (backward) grad angles [3][TR].
template <typename T>
__host__ __device__ constexpr size_t lds_bytes(int tile_res, bool fwd) {
    return sizeof(T) * (size_t)tile_res * (fwd ? 12 : 15);
}

// Coalesced copy of the tile's valid angles into LDS: row k, residues
// [r0, r0+nvalid) -> s_ang[k*TR + j].
template <typename T>
__device__ __forceinline__ void stage_angles(const T* ang_b, int L, int r0, int nvalid, int TR,
                                             T* s_ang) {
    for (int k = 0; k < 3; ++k)
        for (int j = threadIdx.x; j < nvalid; j += blockDim.x)
            s_ang[k * TR + j] = ang_b[(size_t)k * L + r0 + j];
}

// f <- product of this lane's residue transforms (local residues q0..q0+R-1);
// residue r contributes phi_r, psi_r and, if r < n-1, omega_r.
template <typename T, int R>
__device__ __forceinline__ void lane_aggregate(const T* s_ang, int TR, int r0, int n, int q0,
                                               double f[kF]) {
    frame_identity(f);
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int r = r0 + q0 + q;
        if (r < n) {
            step_angle<0>(f, s_ang[q0 + q]);
            step_angle<1>(f, s_ang[TR + q0 + q]);
            if (r < n - 1) step_angle<2>(f, s_ang[2 * TR + q0 + q]);
        }
    }
}

// E <- tile_pfx * wave_aggr[0] * ... * wave_aggr[wave-1] * E, built right to
// left with in-place left products (E enters as the lane's exclusive prefix
// within its wave).
__device__ __forceinline__ void lane_prefix(const double (*s_tot)[kF], int wave,
                                            const double* tile_pfx, double E[kF]) {
    for (int w = wave - 1; w >= 0; --w) frame_lmul(s_tot[w], E);
    if (tile_pfx) frame_lmul(tile_pfx, E);
}

// ------------------------------------------------------------ forward scan
// One tile = TR = 64*W*R residues of chain b.  Steps:
//   1. stage the tile's angles in LDS (coalesced row loads);
//   2. per lane: FP64 sincos and the serial product of its R residues; waves
//      whose first residue is past the chain skip steps 2-4 (identity);
//   3. DPP scan over the wave, LDS over the waves, look-back over tiles;
//   4. per lane: walk its residues from the prefix (sincos recomputed),
//      positions into LDS;
//   5. coalesced store of the tile's coordinates, zeros past the chain.
template <typename T, int R, int MAXT>
__global__ void __launch_bounds__(MAXT) a2b_fwd_scan(ScanArgs a) {
    extern __shared__ __align__(16) unsigned char s_dyn[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nw = blockDim.x >> 6;
    const int TR = blockDim.x * R;
    T* s_ang = reinterpret_cast<T*>(s_dyn);
    T* s_xyz = s_ang + 3 * TR;
    __shared__ double s_tot[4][kF];
    __shared__ double s_pfx[kF];
    __shared__ double s_agg[kF];
    __shared__ int s_tile;
    const T* angles = static_cast<const T*>(a.angles);
    T* coords = static_cast<T*>(a.out);
    const int L = a.L;

    for (int iter = 0;; ++iter) {
        const int tile = next_tile(a, iter, &s_tile);
        if (tile < 0) break;
        const int b = tile % a.B, t = tile / a.B;
        const int n = clamp_length(a, b, t == 0 && tid == 0);
        const int r0 = t * TR;
        const int nvalid = min(max(n - r0, 0), TR);
        T* x_b = coords + (size_t)b * 9 * L;

        if (nvalid > 0) {  // uniform: the tile holds residues of the chain
            stage_angles(angles + (size_t)b * 3 * L, L, r0, nvalid, TR, s_ang);
            __syncthreads();
            const int q0 = tid * R;
            const bool wave_active = wave * 64 * R < nvalid;  // uniform per wave
            double E[kF];
            if (wave_active) {
                lane_aggregate<T, R>(s_ang, TR, r0, n, q0, E);
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
            if (a.lookback) {
                if (tid == 0) {
                    const size_t s0 = (size_t)b * a.ntiles;
                    uint32_t* fl = a.flags + s0;
                    double* pay = a.payload + s0 * kPayload;
                    double* my = pay + (size_t)t * kPayload;
                    double acc[kF];
                    {   // tile aggregate = product of the wave aggregates
#pragma unroll
                        for (int e = 0; e < kF; ++e) acc[e] = s_tot[nw - 1][e];
                        for (int w = nw - 2; w >= 0; --w) frame_lmul(s_tot[w], acc);
#pragma unroll
                        for (int e = 0; e < kF; ++e) {
                            s_agg[e] = acc[e];
                            my[(t == 0 ? kF : 0) + e] = acc[e];
                        }
                    }
                    flag_store(fl + t, t == 0 ? kFlagInclusive : kFlagAggregate);
                    frame_identity(acc);  // becomes this tile's exclusive prefix
                    if (t > 0) {
                        for (int j = t - 1; j >= 0; --j) {
                            const uint32_t fj = wait_flag(fl + j, a.status);
                            if (fj == 0) continue;
                            const double* pj = pay + (size_t)j * kPayload;
                            const bool incl = (fj == kFlagInclusive);
                            frame_lmul(incl ? pj + kF : pj, acc);
                            if (incl) break;
                        }
                        frame_rmul_to(acc, s_agg, my + kF);  // inclusive = prefix * aggregate
                        flag_store(fl + t, kFlagInclusive);
                    }
                    double* ck = a.state + ((size_t)b * a.ntiles + t) * kF;
#pragma unroll
                    for (int e = 0; e < kF; ++e) {
                        s_pfx[e] = acc[e];
                        ck[e] = acc[e];
                    }
                }
                __syncthreads();
            }
            if (wave_active) {
                // Residue r (local j) places CA_r, C_r and N_{r+1}: LDS floats
                // 9j .. 9j+8 stand for global floats 9r+3 .. 9r+11.
                lane_prefix(s_tot, wave, a.lookback ? s_pfx : nullptr, E);
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const int r = r0 + q0 + q;
                    if (r < n) {
                        T* o = s_xyz + 9 * (q0 + q);
                        step_angle<0>(E, s_ang[q0 + q]);
                        o[0] = (T)E[9]; o[1] = (T)E[10]; o[2] = (T)E[11];
                        step_angle<1>(E, s_ang[TR + q0 + q]);
                        o[3] = (T)E[9]; o[4] = (T)E[10]; o[5] = (T)E[11];
                        if (r < n - 1) {
                            step_angle<2>(E, s_ang[2 * TR + q0 + q]);
                            o[6] = (T)E[9]; o[7] = (T)E[10]; o[8] = (T)E[11];
                        }
                    }
                }
            }
            __syncthreads();
        }

        // Coalesced store of global floats [9*r0+3, 9*r0+3+9*TR) within [0, 9L):
        // atoms < 3n from LDS, the rest 0.  Tile 0 also writes N_0 = 0.
        bool bad = false;
        if (t == 0 && tid < 3) x_b[tid] = (T)0;
        const int g0 = 9 * r0 + 3;
        const int gend = min(g0 + 9 * TR, 9 * L);
        const int gvalid = 9 * n;
        for (int g = g0 + tid; g < gend; g += blockDim.x) {
            T v = (T)0;
            if (g < gvalid) {
                v = s_xyz[g - g0];
                bad |= (v != v);
            }
            x_b[g] = v;
        }
        if (bad) set_status(a.status, TPL_A2B_STATUS_NAN);
    }
}

// ----------------------------------------------------------- backward scan
// Same tiling, tickets from the last tile of each chain to the first.  Steps:
//   1. stage angles and the tile's grad_coords in LDS (coalesced);
//   2. lane/wave/tile prefix as in the forward, the tile prefix read from the
//      forward's checkpoint;
//   3. pass 2: positions, and this lane's wrench sum (sum x x g, sum g);
//   4. suffix sums of wrenches over lanes (DPP), waves (LDS) and tiles
//      (reverse look-back);
//   5. pass 3: per atom m, S_m = sum over atoms after m; the gradient is
//      u_m . S_T + (x_m x u_m) . S_F; gradients to LDS;
//   6. coalesced store of the tile's gradients, zeros past the chain.
template <typename T, int R, int MAXT>
__global__ void __launch_bounds__(MAXT) a2b_bwd_scan(ScanArgs a) {
    extern __shared__ __align__(16) unsigned char s_dyn[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nw = blockDim.x >> 6;
    const int TR = blockDim.x * R;
    T* s_ang = reinterpret_cast<T*>(s_dyn);
    T* s_g = s_ang + 3 * TR;
    T* s_out = s_g + 9 * TR;
    __shared__ double s_tot[4][kF];
    __shared__ double s_wsum[4][6];
    __shared__ double s_carry[6];
    __shared__ int s_tile;
    const T* angles = static_cast<const T*>(a.angles);
    const T* gcoords = static_cast<const T*>(a.grad);
    T* gangles = static_cast<T*>(a.out);
    const int L = a.L;

    for (int iter = 0;; ++iter) {
        const int tile = next_tile(a, iter, &s_tile);
        if (tile < 0) break;
        const int b = tile % a.B, t = a.ntiles - 1 - tile / a.B;
        const int n = clamp_length(a, b, t == 0 && tid == 0);
        const int r0 = t * TR;
        const int nvalid = min(max(n - r0, 0), TR);
        T* ga_b = gangles + (size_t)b * 3 * L;

        if (nvalid > 0) {  // uniform
            stage_angles(angles + (size_t)b * 3 * L, L, r0, nvalid, TR, s_ang);
            {
                const T* g_b = gcoords + (size_t)b * 9 * L;
                const int g0 = 9 * r0 + 3;
                const int gend = min(g0 + 9 * TR, 9 * n);
                for (int g = g0 + tid; g < gend; g += blockDim.x) s_g[g - g0] = g_b[g];
            }
            __syncthreads();
            const int q0 = tid * R;
            const bool wave_active = wave * 64 * R < nvalid;
            double E[kF];
            if (wave_active) {
                lane_aggregate<T, R>(s_ang, TR, r0, n, q0, E);
                wave_scan_frames(E);
                if (lane == 63) {
#pragma unroll
                    for (int e = 0; e < kF; ++e) s_tot[wave][e] = E[e];
                }
                wave_shift_frames(E);
            }
            __syncthreads();
            double W[6] = {0, 0, 0, 0, 0, 0};  // (sum x x g, sum g) of this lane
            if (wave_active) {
                lane_prefix(s_tot, wave,
                            a.lookback ? a.state + ((size_t)b * a.ntiles + t) * kF : nullptr, E);
                double X[kF];
#pragma unroll
                for (int e = 0; e < kF; ++e) X[e] = E[e];
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const int r = r0 + q0 + q;
                    if (r < n) {
                        const int na = (r < n - 1) ? 3 : 2;
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            if (k < na) {
                                if (k == 0) step_angle<0>(X, s_ang[q0 + q]);
                                if (k == 1) step_angle<1>(X, s_ang[TR + q0 + q]);
                                if (k == 2) step_angle<2>(X, s_ang[2 * TR + q0 + q]);
                                const T* gp = s_g + 9 * (q0 + q) + 3 * k;
                                const double g0 = gp[0], g1 = gp[1], g2 = gp[2];
                                const double x0 = X[9], x1 = X[10], x2 = X[11];
                                W[0] += x1 * g2 - x2 * g1;
                                W[1] += x2 * g0 - x0 * g2;
                                W[2] += x0 * g1 - x1 * g0;
                                W[3] += g0; W[4] += g1; W[5] += g2;
                            }
                        }
                    }
                }
            }
            // Suffix sums of wrenches: lanes, waves and tiles after this one.
            double C[6];
#pragma unroll
            for (int e = 0; e < 6; ++e) C[e] = W[e];
            if (wave_active) {
                wave_scan_sum<6>(C);  // inclusive prefix over lanes
#pragma unroll
                for (int e = 0; e < 6; ++e) {
                    const double tot = read_lane(C[e], 63);
                    if (lane == 0) s_wsum[wave][e] = tot;
                    C[e] = tot - C[e];  // sum over later lanes
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
            if (a.lookback) {
                if (tid == 0) {
                    double tot[6] = {0, 0, 0, 0, 0, 0}, carry[6] = {0, 0, 0, 0, 0, 0};
                    for (int w = 0; w < nw; ++w) {
#pragma unroll
                        for (int e = 0; e < 6; ++e) tot[e] += s_wsum[w][e];
                    }
                    const int ntv = (n + TR - 1) / TR;  // tiles holding residues
                    const size_t s0 = (size_t)b * a.ntiles;
                    uint32_t* fl = a.flags + s0;
                    double* pay = a.payload + s0 * kPayload;
                    double* my = pay + (size_t)t * kPayload;
                    if (t == ntv - 1) {
#pragma unroll
                        for (int e = 0; e < 6; ++e) my[kF + e] = tot[e];
                        flag_store(fl + t, kFlagInclusive);
                    } else {
#pragma unroll
                        for (int e = 0; e < 6; ++e) my[e] = tot[e];
                        flag_store(fl + t, kFlagAggregate);
                        for (int j = t + 1; j < ntv; ++j) {
                            const uint32_t fj = wait_flag(fl + j, a.status);
                            if (fj == 0) continue;
                            const double* pj = pay + (size_t)j * kPayload;
                            const bool incl = (fj == kFlagInclusive);
#pragma unroll
                            for (int e = 0; e < 6; ++e) carry[e] += incl ? pj[kF + e] : pj[e];
                            if (incl) break;
                        }
#pragma unroll
                        for (int e = 0; e < 6; ++e) my[kF + e] = carry[e] + tot[e];
                        flag_store(fl + t, kFlagInclusive);
                    }
#pragma unroll
                    for (int e = 0; e < 6; ++e) s_carry[e] = carry[e];
                }
                __syncthreads();
#pragma unroll
                for (int e = 0; e < 6; ++e) C[e] += s_carry[e];
            }
            if (wave_active) {
                // Pass 3: S_m = sum over atoms after m; gradients into LDS.
                double S[6];
#pragma unroll
                for (int e = 0; e < 6; ++e) S[e] = C[e] + W[e];
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const int r = r0 + q0 + q;
                    if (r < n) {
                        double gr[3] = {0.0, 0.0, 0.0};
                        const int na = (r < n - 1) ? 3 : 2;
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            if (k < na) {
                                if (k == 0) step_angle<0>(E, s_ang[q0 + q]);
                                if (k == 1) step_angle<1>(E, s_ang[TR + q0 + q]);
                                if (k == 2) step_angle<2>(E, s_ang[2 * TR + q0 + q]);
                                const T* gp = s_g + 9 * (q0 + q) + 3 * k;
                                const double g0 = gp[0], g1 = gp[1], g2 = gp[2];
                                const double x0 = E[9], x1 = E[10], x2 = E[11];
                                const double u0 = E[0], u1 = E[1], u2 = E[2];
                                S[0] -= x1 * g2 - x2 * g1;
                                S[1] -= x2 * g0 - x0 * g2;
                                S[2] -= x0 * g1 - x1 * g0;
                                S[3] -= g0; S[4] -= g1; S[5] -= g2;
                                const double w0 = x1 * u2 - x2 * u1;
                                const double w1 = x2 * u0 - x0 * u2;
                                const double w2 = x0 * u1 - x1 * u0;
                                gr[k] = u0 * S[0] + u1 * S[1] + u2 * S[2] + w0 * S[3] +
                                        w1 * S[4] + w2 * S[5];
                            }
                        }
                        if (r == n - 1) gr[1] = 0.0;  // psi of the last residue moves no atom
                        s_out[q0 + q] = (T)gr[0];
                        s_out[TR + q0 + q] = (T)gr[1];
                        s_out[2 * TR + q0 + q] = (T)gr[2];
                    }
                }
            }
            __syncthreads();
        }

        // Coalesced store of the tile's gradients, zeros past the chain.
        bool bad = false;
        const int jend = min(TR, L - r0);
        for (int k = 0; k < 3; ++k) {
            T* dst = ga_b + (size_t)k * L + r0;
            for (int j = tid; j < jend; j += blockDim.x) {
                T v = (T)0;
                if (j < nvalid) {
                    v = s_out[k * TR + j];
                    bad |= (v != v);
                }
                dst[j] = v;
            }
        }
        if (bad) set_status(a.status, TPL_A2B_STATUS_NAN);
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

// Largest tile whose LDS staging (see lds_bytes) stays within 64 KiB.
constexpr int kMaxTileF32 = 1024, kMaxTileF64 = 512;

bool fits_dtype(int dtype, const tpl_a2b_config* c) {
    const int tr = 64 * c->waves * c->res_per_lane;
    return tr <= (dtype == TPL_A2B_F32 ? kMaxTileF32 : kMaxTileF64);
}

int tiles_for(int L, const tpl_a2b_config* c) {
    const int tr = 64 * c->waves * c->res_per_lane;
    return L <= 0 ? 0 : (L + tr - 1) / tr;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct SyncLayout {
    size_t flags_off, ticket_off, words, payload_off, total;
};

SyncLayout sync_layout(int B, int nt) {
    SyncLayout s{};
    const size_t tiles = (size_t)B * nt;
    s.flags_off = 0;
    s.ticket_off = tiles * 4;
    s.words = tiles + 1;  // flags + ticket, all zeroed per call
    s.payload_off = align256(s.words * 4);
    s.total = s.payload_off + tiles * kPayload * sizeof(double);
    return s;
}

using KernelFn = void (*)(ScanArgs);

constexpr int kMaxThreads = 256;  // up to 4 waves per workgroup

template <typename T, int R>
KernelFn pick_kernel(bool fwd) {
    return fwd ? a2b_fwd_scan<T, R, kMaxThreads> : a2b_bwd_scan<T, R, kMaxThreads>;
}

template <typename T>
KernelFn pick_kernel(bool fwd, const tpl_a2b_config* c) {
    switch (c->res_per_lane) {
        case 1: return pick_kernel<T, 1>(fwd);
        case 2: return pick_kernel<T, 2>(fwd);
        case 4: return pick_kernel<T, 4>(fwd);
        default: return pick_kernel<T, 8>(fwd);
    }
}

int check_common(int dtype, int B, int L) {
    if (dtype != TPL_A2B_F32 && dtype != TPL_A2B_F64) return TPL_A2B_ERR_ARG;
    if (B < 0 || L < 0) return TPL_A2B_ERR_ARG;
    if ((int64_t)B * 9 * (int64_t)L > (int64_t)INT32_MAX * 4) return TPL_A2B_ERR_SIZE;
    return TPL_A2B_OK;
}

int launch_scan(bool fwd, int dtype, ScanArgs args, void* sync, const tpl_a2b_config* cfg,
                hipStream_t stream) {
    const int nt = args.ntiles;
    const int total = args.B * nt;
    args.lookback = nt > 1 ? 1 : 0;
    if (args.lookback) {
        const SyncLayout sl = sync_layout(args.B, nt);
        char* base = static_cast<char*>(sync);
        args.flags = reinterpret_cast<uint32_t*>(base + sl.flags_off);
        args.ticket = reinterpret_cast<uint32_t*>(base + sl.ticket_off);
        args.payload = reinterpret_cast<double*>(base + sl.payload_off);
        const int zb = (int)std::min<size_t>((sl.words + 255) / 256, 1024);
        hipLaunchKernelGGL(a2b_zero_words, dim3(zb), dim3(256), 0, stream, args.flags,
                           (int)sl.words);
        if (hipGetLastError() != hipSuccess) return TPL_A2B_ERR_LAUNCH;
    }
    int grid = total;
    if (cfg->max_wgs > 0) grid = std::min(grid, cfg->max_wgs);
    const int threads = 64 * cfg->waves;
    const int tr = threads * cfg->res_per_lane;
    const bool f32 = dtype == TPL_A2B_F32;
    KernelFn k = f32 ? pick_kernel<float>(fwd, cfg) : pick_kernel<double>(fwd, cfg);
    const size_t lds = f32 ? lds_bytes<float>(tr, fwd) : lds_bytes<double>(tr, fwd);
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

size_t tpl_a2b_state_bytes(int batch, int angles_stride, const tpl_a2b_config* cfg) {
    if (!valid_config(cfg) || batch <= 0 || angles_stride <= 0) return 0;
    const int nt = tiles_for(angles_stride, cfg);
    return nt > 1 ? (size_t)batch * nt * kF * sizeof(double) : 0;
}

size_t tpl_a2b_sync_bytes(int batch, int angles_stride, const tpl_a2b_config* cfg) {
    if (!valid_config(cfg) || batch <= 0 || angles_stride <= 0) return 0;
    const int nt = tiles_for(angles_stride, cfg);
    return nt > 1 ? sync_layout(batch, nt).total : 0;
}

int tpl_a2b_forward(int dtype, const void* angles, void* coords, const int32_t* lengths,
                    void* state, void* sync, uint32_t* status, int batch, int angles_stride,
                    const tpl_a2b_config* cfg, tpl_stream_t stream) {
    int rc = check_common(dtype, batch, angles_stride);
    if (rc) return rc;
    if (!valid_config(cfg) || !fits_dtype(dtype, cfg)) return TPL_A2B_ERR_CONFIG;
    if (batch == 0 || angles_stride == 0) return TPL_A2B_OK;
    if (!angles || !coords || !lengths) return TPL_A2B_ERR_ARG;
    const int nt = tiles_for(angles_stride, cfg);
    if (nt > 1 && (!state || !sync)) return TPL_A2B_ERR_ARG;
    ScanArgs a{};
    a.angles = angles;
    a.out = coords;
    a.lengths = lengths;
    a.state = static_cast<double*>(state);
    a.status = status;
    a.B = batch;
    a.L = angles_stride;
    a.ntiles = nt;
    return launch_scan(true, dtype, a, sync, cfg, static_cast<hipStream_t>(stream));
}

int tpl_a2b_backward(int dtype, void* grad_angles, const void* grad_coords, const void* angles,
                     const int32_t* lengths, const void* state, void* sync, uint32_t* status,
                     int batch, int angles_stride, const tpl_a2b_config* cfg,
                     tpl_stream_t stream) {
    int rc = check_common(dtype, batch, angles_stride);
    if (rc) return rc;
    if (!valid_config(cfg) || !fits_dtype(dtype, cfg)) return TPL_A2B_ERR_CONFIG;
    if (batch == 0 || angles_stride == 0) return TPL_A2B_OK;
    if (!grad_angles || !grad_coords || !angles || !lengths) return TPL_A2B_ERR_ARG;
    const int nt = tiles_for(angles_stride, cfg);
    if (nt > 1 && (!state || !sync)) return TPL_A2B_ERR_ARG;
    ScanArgs a{};
    a.angles = angles;
    a.out = grad_angles;
    a.grad = grad_coords;
    a.lengths = lengths;
    a.state = const_cast<double*>(static_cast<const double*>(state));
    a.status = status;
    a.B = batch;
    a.L = angles_stride;
    a.ntiles = nt;
    return launch_scan(false, dtype, a, sync, cfg, static_cast<hipStream_t>(stream));
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
                rc = tpl_a2b_backward(dtype, grad_angles, grad_coords, angles, lengths, state,
                                      sync, nullptr, batch, angles_stride, cf, stream);
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
