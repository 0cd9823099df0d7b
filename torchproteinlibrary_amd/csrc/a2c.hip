// Full-atom protein builder for MI355X (gfx950): FullAtomModel.Angles2Coords
// forward and backward.  The C ABI is include/tpl_a2c.h.
//
// The reference builds a tree of heap-allocated 4x4 transforms per structure
// on the CPU and walks it recursively (Layers/FullAtomModel/cConformation.cpp:
// 29-158); its backward visits, for every angle, every atom below it
// (:160-195), O(atoms x tree depth) with a chain's depth growing with its
// length.  Here one workgroup takes one structure:
//
//   1. each thread takes a contiguous run of residues and multiplies their
//      residue steps S_i = T_CA(phi_i) T_C(psi_i) T_N(omega_i) (FP64 3x4), the
//      transform from residue i's N frame to residue i+1's;
//   2. a workgroup scan of those products (DPP within each wave, the waves'
//      totals through LDS) gives each thread the N frame of its first
//      residue, and a scan of the atom counts gives its first atom;
//   3. each thread walks its residues: N, CA, the side chain, C, writing atoms.
//
// The backward uses dL/dalpha = u . (sum_{x in subtree} x cross g
// - p cross sum_{x in subtree} g) for a node with x axis u and origin p (the
// reference's g . (F x) with F = M_parent dT M^-1, cConformation.cpp:148,
// 171-184, reduces to this since F x = u cross (x - p)).  u and p follow from
// the forward's coordinates, so the backward reads those and needs no frames.
// The subtree of a backbone node is the rest of its residue plus every later
// residue: a workgroup scan of per-run (sum g, sum x cross g), with runs in
// reverse thread order so that it is a prefix scan, supplies the downstream
// part, and each thread walks its residues once more.  For batches that fill the
// GPU a second backward walks each structure in tiles of consecutive atoms
// instead (a2c_bwd_tiled), with coalesced loads and each atom read once.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <vector>

#include "a2c_table.h"
#include "sincos_tab.h"
#include "tpl_a2c.h"

namespace {

using namespace a2c;

// Workgroup sizes: 256 threads, or 512 for long structures (see use_large_fwd
// and use_large_bwd), where a thread then walks half as many residues.  LDS:
// forward 13.6 / 14.0 KB (the tables), run backward 0.8 / 1.0 KB.
constexpr int kThreadsSmall = 256, kThreadsLarge = 512;

__constant__ Residue kRes[kTypes] = A2C_RESIDUES;
__constant__ Side kSide[kTypes][kMaxSide] = A2C_SIDES;
__constant__ Atom kAtom[kTypes][kMaxAtoms] = A2C_ATOMS;

// Workgroup-shared tables, filled once per workgroup from the constant ones:
// local atom positions, and per side-chain node cos/sin of beta and of a
// constant alpha, d, and the packed (row, first atom, count).
struct Tables {
    double loc[kTypes * kMaxAtoms][3];
    double side[kTypes * kMaxSide][5];
    int side_i[kTypes * kMaxSide];
    int res_i[kTypes];  // natoms | nca << 8 | nside << 16
    double bb[3][2];    // cos/sin of beta for the CA, C and N transforms
    double tab[128];    // the sincos table of sincos_tab.h
};

// The table sincos (sincos_tab.h) over the workgroup's LDS copy of its table.
struct TabSinCos {
    const double* tab;
    __device__ __forceinline__ void operator()(double x, double* s, double* c) const {
        sincos_tab(x, tab, s, c);
    }
};

// The sines and cosines of the constant angles come from the table sincos
// (within 2.2e-16 of the math library's): the library's FP64 sincos took
// 4-6 us of every workgroup's start.
template <int NT>
__device__ void fill_tables(Tables& tb) {
    for (int k = threadIdx.x; k < 128; k += NT) tb.tab[k] = kTabCS[k];
    __syncthreads();  // tb.tab, for the rest
    const TabSinCos sc{tb.tab};
    for (int k = threadIdx.x; k < kTypes * kMaxAtoms; k += NT) {
        const Atom at = kAtom[k / kMaxAtoms][k % kMaxAtoms];
        local_position_with(at, tb.loc[k], sc);
    }
    for (int k = threadIdx.x; k < kTypes * kMaxSide; k += NT) {
        const Side s = kSide[k / kMaxSide][k % kMaxSide];
        sc(s.beta, &tb.side[k][1], &tb.side[k][0]);
        sc(s.alpha, &tb.side[k][3], &tb.side[k][2]);
        tb.side[k][4] = s.d;
        tb.side_i[k] = (s.row & 0xff) | (s.first << 8) | (s.count << 16);
    }
    if (threadIdx.x < kTypes) {
        const Residue r = kRes[threadIdx.x];
        tb.res_i[threadIdx.x] = r.natoms | (r.nca << 8) | (r.nside << 16);
    }
    if (threadIdx.x < 3) {
        const double beta = threadIdx.x == 0 ? geo::C_N_CA : (threadIdx.x == 1 ? geo::N_CA_C : geo::CA_C_N);
        sc(beta, &tb.bb[threadIdx.x][1], &tb.bb[threadIdx.x][0]);
    }
}

// The tables are the same for every workgroup and call: a one-workgroup
// kernel computes them once per device into g_tab, and each forward
// workgroup copies that (1.3 us less per workgroup than computing them:
// 32 x 350 residues 17.5 -> 16.2 us, 1024 x 30 15.2 -> 13.4 us).
__device__ Tables g_tab;

__global__ void __launch_bounds__(256) a2c_init_tables() {
    __shared__ Tables tb;
    fill_tables<256>(tb);
    __syncthreads();
    const double* src = reinterpret_cast<const double*>(&tb);
    double* dst = reinterpret_cast<double*>(&g_tab);
    for (int k = threadIdx.x; k < (int)(sizeof(Tables) / 8); k += 256) dst[k] = src[k];
}

template <int NT>
__device__ __forceinline__ void copy_tables(Tables& tb) {
    const double* src = reinterpret_cast<const double*>(&g_tab);
    double* dst = reinterpret_cast<double*>(&tb);
    for (int k = threadIdx.x; k < (int)(sizeof(Tables) / 8); k += NT) dst[k] = src[k];
}

// m (3x4, row-major) <- m * Ry(beta) Tx(d) Rx(alpha), with cb/sb of beta and
// ca/sa of alpha.  Ry(b) Tx(d) Rx(a) = [[cb, sb sa, sb ca, d cb],
// [0, ca, -sa, 0], [-sb, cb sa, cb ca, -d sb]].
__device__ __forceinline__ void step(double m[12], double cb, double sb, double ca, double sa,
                                     double d) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double a0 = m[4 * i], a1 = m[4 * i + 1], a2 = m[4 * i + 2];
        const double n0 = a0 * cb - a2 * sb;
        const double t = a0 * sb + a2 * cb;
        m[4 * i] = n0;
        m[4 * i + 1] = t * sa + a1 * ca;
        m[4 * i + 2] = t * ca - a1 * sa;
        m[4 * i + 3] += d * n0;
    }
}

// c = a * b for 3x4 rigid transforms.
__device__ __forceinline__ void compose(const double a[12], const double b[12], double c[12]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
            c[4 * i + j] = a[4 * i] * b[j] + a[4 * i + 1] * b[4 + j] + a[4 * i + 2] * b[8 + j] +
                           (j == 3 ? a[4 * i + 3] : 0.0);
    }
}

__device__ __forceinline__ void identity(double m[12]) {
#pragma unroll
    for (int k = 0; k < 12; ++k) m[k] = (k == 0 || k == 5 || k == 10) ? 1.0 : 0.0;
}

__device__ __forceinline__ void apply(const double m[12], const double v[3], double x[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) x[i] = m[4 * i] * v[0] + m[4 * i + 1] * v[1] + m[4 * i + 2] * v[2] + m[4 * i + 3];
}

// ---- wave-level scans (DPP; all 64 lanes of the wave active) ----
// GFX9 DPP controls: row shifts, row broadcasts, wave shift right by one.
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143, kWaveShr1 = 0x138;

// DPP move of a double; lanes that receive nothing (or rows masked off)
// keep `old`.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_double(double v, double old) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(v), CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(v), CTRL, ROW_MASK, 0xf, false);
    return __hiloint2double(hi, lo);
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp_int(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, 0xf, false);
}

__device__ __forceinline__ double ident12(int k) {
    return (k == 0 || k == 5 || k == 10) ? 1.0 : 0.0;
}

// One Hillis-Steele step of the ordered product: m <- (source lane's m) * m,
// with the identity where a lane receives nothing.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void frame_scan_step(double m[12]) {
    double g[12], r[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) g[k] = dpp_double<CTRL, ROW_MASK>(m[k], ident12(k));
    compose(g, m, r);
#pragma unroll
    for (int k = 0; k < 12; ++k) m[k] = r[k];
}

// Inclusive ordered product over the wave's lanes (lane l ends with
// m_0 * ... * m_l).
__device__ __forceinline__ void wave_scan_frames(double m[12]) {
    frame_scan_step<kRowShr1, 0xf>(m);
    frame_scan_step<kRowShr2, 0xf>(m);
    frame_scan_step<kRowShr4, 0xf>(m);
    frame_scan_step<kRowShr8, 0xf>(m);
    frame_scan_step<kRowBcast15, 0xa>(m);
    frame_scan_step<kRowBcast31, 0xc>(m);
}

// Inclusive prefix sums over the wave's lanes.
__device__ __forceinline__ double wave_scan_double(double v) {
    v += dpp_double<kRowShr1, 0xf>(v, 0.0);
    v += dpp_double<kRowShr2, 0xf>(v, 0.0);
    v += dpp_double<kRowShr4, 0xf>(v, 0.0);
    v += dpp_double<kRowShr8, 0xf>(v, 0.0);
    v += dpp_double<kRowBcast15, 0xa>(v, 0.0);
    v += dpp_double<kRowBcast31, 0xc>(v, 0.0);
    return v;
}
// Inclusive prefix sums within each row of 16 lanes.
__device__ __forceinline__ double row_scan_double(double v) {
    v += dpp_double<kRowShr1, 0xf>(v, 0.0);
    v += dpp_double<kRowShr2, 0xf>(v, 0.0);
    v += dpp_double<kRowShr4, 0xf>(v, 0.0);
    v += dpp_double<kRowShr8, 0xf>(v, 0.0);
    return v;
}
// The value of lane `lane` (wave-uniform) of a double.
__device__ __forceinline__ double readlane_double(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ int wave_scan_int(int v) {
    v += dpp_int<kRowShr1, 0xf>(v);
    v += dpp_int<kRowShr2, 0xf>(v);
    v += dpp_int<kRowShr4, 0xf>(v);
    v += dpp_int<kRowShr8, 0xf>(v);
    v += dpp_int<kRowBcast15, 0xa>(v);
    v += dpp_int<kRowBcast31, 0xc>(v);
    return v;
}

// The structure's length: the first 0 byte of its sequence row, or
// seq_stride (s_len enters as seq_stride).  Each thread reads 4 consecutive
// bytes per round, the loads issued together.
template <int NT>
__device__ __forceinline__ void find_length(const uint8_t* seq, int seq_stride, int* s_len) {
    constexpr int kPer = 4;
    for (int i0 = (int)threadIdx.x * kPer; i0 < seq_stride; i0 += NT * kPer) {
        uint8_t c[kPer];
#pragma unroll
        for (int e = 0; e < kPer; ++e) c[e] = i0 + e < seq_stride ? seq[i0 + e] : (uint8_t)1;
        int first = -1;
#pragma unroll
        for (int e = kPer - 1; e >= 0; --e)
            if (c[e] == 0) first = i0 + e;
        if (first >= 0) atomicMin(s_len, first);
    }
}

template <typename T>
__device__ __forceinline__ double angle(const T* ang, int row, int stride, int i) {
    return (double)ang[(size_t)row * stride + i];
}

// The 8 angles of residue i, loaded together at the top of the residue so
// that their latencies overlap (one global-load latency per residue walk
// instead of one per transform).
template <typename T>
struct ResAngles {
    T v[8];
    __device__ __forceinline__ void load(const T* ang, int stride, int i) {
#pragma unroll
        for (int r = 0; r < 8; ++r) v[r] = ang[(size_t)r * stride + i];
    }
};

// The side-chain angles xi1..xi5 of a residue in the order its side-chain
// nodes bind them: every template binds rows 3, 4, ... consecutively (the
// table test checks this), so each bound node takes the next one.  Shifting
// keeps the values in registers where an index computed at run time would not.
template <typename T>
struct SideRows {
    T v[5];
    __device__ __forceinline__ explicit SideRows(const ResAngles<T>& a) {
#pragma unroll
        for (int k = 0; k < 5; ++k) v[k] = a.v[3 + k];
    }
    __device__ __forceinline__ double pop() {
        const T x = v[0];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = v[k + 1];
        return (double)x;
    }
};

// The residue step S_i applied to m: m <- m T_CA(phi) T_C(psi) T_N(omega).
template <typename T>
__device__ __forceinline__ void residue_step(double m[12], const Tables& tb, const T* ang, int stride,
                                             int i) {
    const double phi = angle(ang, 0, stride, i), psi = angle(ang, 1, stride, i),
                 omega = angle(ang, 2, stride, i);
    double s, c;
    sincos_tab(phi, tb.tab, &s, &c);
    step(m, tb.bb[0][0], tb.bb[0][1], c, s, geo::R_N_CA);
    sincos_tab(psi, tb.tab, &s, &c);
    step(m, tb.bb[1][0], tb.bb[1][1], c, s, geo::R_CA_C);
    sincos_tab(omega, tb.tab, &s, &c);
    step(m, tb.bb[2][0], tb.bb[2][1], c, s, geo::R_C_N);
}

// The structure this workgroup builds: order[blockIdx.x], or blockIdx.x
// without an order.  An index outside the batch is reported
// (TPL_A2C_STATUS_ORDER) and the workgroup does nothing (-1).
__device__ __forceinline__ int structure_of(const int* order, uint32_t* status) {
    if (!order) return (int)blockIdx.x;
    const int b = __builtin_amdgcn_readfirstlane(order[blockIdx.x]);
    if (b >= 0 && b < (int)gridDim.x) return b;
    if (threadIdx.x == 0 && status)
        __hip_atomic_fetch_or(status, TPL_A2C_STATUS_ORDER, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return -1;
}

// Per structure: the sequence length (first 0 byte), the residues' validity,
// this thread's run [r0, r1), and the workgroup scans giving the N frame of
// residue r0 (frame) and its first atom (atom0).  Returns false (for every
// thread of the workgroup) if the structure is flagged.
struct Run {
    int L, r0, r1, atom0, total;
};

template <int NT, typename T>
__device__ bool prepare(const uint8_t* seq, int seq_stride, const T* ang, int ang_stride,
                        int atoms_stride, uint32_t* status, const Tables& tb, double (*s_m)[NT / 64],
                        int* s_c, int* s_flag, Run& run, double frame[12], ResAngles<T>& av0,
                        uint8_t& code0) {
    const int tid = threadIdx.x;
    if (tid == 0) {
        s_flag[0] = seq_stride;  // length
        s_flag[1] = 0;           // status bits
    }
    __syncthreads();
    find_length<NT>(seq, seq_stride, &s_flag[0]);
    __syncthreads();
    int L = s_flag[0];
    if (L > ang_stride) {
        if (tid == 0) atomicOr(&s_flag[1], (int)TPL_A2C_STATUS_LENGTH);
        L = 0;
    }
    const int chunk = (L + NT - 1) / NT;
    const int r0 = min(L, tid * chunk), r1 = min(L, r0 + chunk);
    // The walk's first residue: its angles and letter are loaded now, so
    // that they arrive while the workgroup scans.
    if (r0 < r1) {
        av0.load(ang, ang_stride, r0);
        code0 = seq[r0];
    }
    double P[12];
    identity(P);
    int cnt = 0;
    bool bad = false;
    for (int i = r0; i < r1; ++i) {
        const int t = residue_type(seq[i]);
        if (t < 0) {
            bad = true;
            continue;
        }
        cnt += tb.res_i[t] & 0xff;
        residue_step(P, tb, ang, ang_stride, i);
    }
    if (bad) atomicOr(&s_flag[1], (int)TPL_A2C_STATUS_SEQ);
    // Ordered product of the runs' steps and sum of their atom counts over
    // the workgroup: inclusive within each wave by DPP, then across waves
    // through LDS (the waves' totals; each wave composes those before it).
    // frame = product of the earlier runs' steps, atom0 = their atoms.
    constexpr int NW = NT / 64;
    const int wv = tid >> 6;
    int cinc = cnt;
    if ((tid & ~63) * chunk < L) {  // a wave whose runs are all past the structure holds identities and zeros
        cinc = wave_scan_int(cnt);
        wave_scan_frames(P);
    }
    if ((tid & 63) == 63) {
#pragma unroll
        for (int k = 0; k < 12; ++k) s_m[k][wv] = P[k];
        s_c[wv] = cinc;
    }
    double E[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) E[k] = dpp_double<kWaveShr1, 0xf>(P[k], ident12(k));
    __syncthreads();
    double W[12];
    identity(W);
    int before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        if (w < wv) {
            double Tw[12], R2[12];
#pragma unroll
            for (int k = 0; k < 12; ++k) Tw[k] = s_m[k][w];
            compose(W, Tw, R2);
#pragma unroll
            for (int k = 0; k < 12; ++k) W[k] = R2[k];
            before += s_c[w];
        }
        total += s_c[w];
    }
    compose(W, E, frame);
    run.atom0 = before + cinc - cnt;
    run.total = total;
    if (run.total > atoms_stride && tid == 0) atomicOr(&s_flag[1], (int)TPL_A2C_STATUS_ATOMS);
    __syncthreads();
    const int flags = s_flag[1];
    if (flags && tid == 0 && status) __hip_atomic_fetch_or(status, (uint32_t)flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    run.L = L;
    run.r0 = r0;
    run.r1 = r1;
    if (flags) {
        run.L = run.r0 = run.r1 = run.total = 0;
        return false;
    }
    return true;
}

// At 256 threads the forward is held to four waves per SIMD (<= 128 VGPRs,
// a few dwords spilled) so that four workgroups fit on a CU, as LDS allows:
// 85 -> 71 us at mixed lengths 30-1000, B=1024, against three workgroups.
template <int NT, typename T>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT == 256 ? 4 : 1, 8))) a2c_fwd(const uint8_t* __restrict__ seqs, int seq_stride,
                                                    const T* __restrict__ angles, int ang_stride,
                                                    T* __restrict__ coords, uint32_t* __restrict__ res_names,
                                                    uint32_t* __restrict__ atom_names, int atoms_stride,
                                                    uint32_t* status, const int* __restrict__ order) {
    __shared__ Tables tb;
    __shared__ double s_m[12][NT / 64];  // the waves' products
    __shared__ int s_c[NT / 64];          // the waves' atom counts
    __shared__ int s_flag[2];
    const int b = structure_of(order, status);
    if (b < 0) return;
    const uint8_t* seq = seqs + (size_t)b * seq_stride;
    const T* ang = angles + (size_t)b * 8 * ang_stride;
    T* out = coords + (size_t)b * 3 * atoms_stride;
    copy_tables<NT>(tb);
    __syncthreads();
    Run run;
    double G[12];
    ResAngles<T> av;
    uint8_t code = 0;
    prepare<NT>(seq, seq_stride, ang, ang_stride, atoms_stride, status, tb, s_m, s_c, s_flag, run, G, av, code);
    int a = run.atom0;
    uint32_t* rn = res_names ? res_names + (size_t)b * atoms_stride : nullptr;
    uint32_t* an = atom_names ? atom_names + (size_t)b * atoms_stride : nullptr;
    for (int i = run.r0; i < run.r1; ++i) {
        // the next residue's angles and letter go out before this one's work
        ResAngles<T> nav;
        uint8_t ncode = 0;
        if (i + 1 < run.r1) {
            nav.load(ang, ang_stride, i + 1);
            ncode = seq[i + 1];
        }
        const int t = residue_type(code);
        const int ri = tb.res_i[t];
        const int natoms = ri & 0xff, nca = (ri >> 8) & 0xff, nside = ri >> 16;
        const double* loc = tb.loc[t * kMaxAtoms];
        double x[3];
        auto put = [&](int slot, const double m[12]) {
            apply(m, loc + 3 * slot, x);
            T* o = out + 3 * (size_t)(a + slot);
            o[0] = (T)x[0];
            o[1] = (T)x[1];
            o[2] = (T)x[2];
        };
        put(0, G);
        double s, c, F[12], CA[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) CA[k] = G[k];
        sincos_tab((double)av.v[0], tb.tab, &s, &c);
        step(CA, tb.bb[0][0], tb.bb[0][1], c, s, geo::R_N_CA);
        for (int k = 1; k <= nca; ++k) put(k, CA);
#pragma unroll
        for (int k = 0; k < 12; ++k) F[k] = CA[k];
        SideRows<T> rows(av);
        for (int n = 0; n < nside; ++n) {
            const int q = t * kMaxSide + n;
            const int si = tb.side_i[q];
            const int row = (int)(int8_t)(si & 0xff), first = (si >> 8) & 0xff, count = si >> 16;
            double ca = tb.side[q][2], sa = tb.side[q][3];
            if (row >= 0) sincos_tab(rows.pop(), tb.tab, &sa, &ca);
            step(F, tb.side[q][0], tb.side[q][1], ca, sa, tb.side[q][4]);
            for (int k = first; k < first + count; ++k) put(k, F);
        }
        sincos_tab((double)av.v[1], tb.tab, &s, &c);
        step(CA, tb.bb[1][0], tb.bb[1][1], c, s, geo::R_CA_C);  // CA now holds the C frame
        put(natoms - 2, CA);
        put(natoms - 1, CA);
        if (rn) {
            const uint32_t r3 = *reinterpret_cast<const uint32_t*>(kRes[t].name3);
            for (int k = 0; k < natoms; ++k) {
                rn[a + k] = r3;
                an[a + k] = *reinterpret_cast<const uint32_t*>(kAtom[t][k].name);
            }
        }
        sincos_tab((double)av.v[2], tb.tab, &s, &c);
        step(CA, tb.bb[2][0], tb.bb[2][1], c, s, geo::R_C_N);
#pragma unroll
        for (int k = 0; k < 12; ++k) G[k] = CA[k];
        a += natoms;
        av = nav;
        code = ncode;
    }
    for (int k = 3 * run.total + threadIdx.x; k < 3 * atoms_stride; k += NT) out[k] = (T)0;
    if (rn)
        for (int k = run.total + threadIdx.x; k < atoms_stride; k += NT) rn[k] = an[k] = 0u;
}

struct Sums {
    double g[3], xg[3];
};

__device__ __forceinline__ void sums_zero(Sums& s) {
#pragma unroll
    for (int c = 0; c < 3; ++c) s.g[c] = s.xg[c] = 0.0;
}

__device__ __forceinline__ void sums_add(Sums& s, const double x[3], const double g[3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) s.g[c] += g[c];
    s.xg[0] += x[1] * g[2] - x[2] * g[1];
    s.xg[1] += x[2] * g[0] - x[0] * g[2];
    s.xg[2] += x[0] * g[1] - x[1] * g[0];
}

// Backward from the forward's coordinates.  A node bound to an angle turns
// its subtree about its x axis through its origin; the origin is the node's
// first atom (every group's first atom sits at the group origin), and the x
// axis is the unit bond vector from the parent node's origin (the transform
// Ry Tx(d) Rx moves the origin by d along Ry's x axis, which Rx keeps), so
// the gradient needs no frames: dL/dalpha = u . (sum x cross g - p cross
// sum g) over the subtree, with u = (p - q) / |p - q| for parent origin q.

// Per structure, without frames: length, validity, this thread's run and its
// first atom (workgroup scan of the atom counts).  Flags as in prepare().
// With s_ra, also each residue's first atom | type << 24 (visible to the
// workgroup after its next barrier).
template <int NT>
__device__ void count_run(const uint8_t* seq, int seq_stride, int ang_stride, int atoms_stride,
                          uint32_t* status, const int* res_i, int* s_c, int* s_flag, Run& run,
                          int* s_ra = nullptr) {
    const int tid = threadIdx.x;
    if (tid == 0) {
        s_flag[0] = seq_stride;
        s_flag[1] = 0;
    }
    __syncthreads();
    find_length<NT>(seq, seq_stride, &s_flag[0]);
    __syncthreads();
    int L = s_flag[0];
    if (L > ang_stride) {
        if (tid == 0) atomicOr(&s_flag[1], (int)TPL_A2C_STATUS_LENGTH);
        L = 0;
    }
    // Runs go to threads in reverse order (run q = NT - 1 - tid), so that
    // sums over a run and every later run are prefix sums over threads.
    const int chunk = (L + NT - 1) / NT;
    const int q = NT - 1 - tid;
    const int r0 = min(L, q * chunk), r1 = min(L, r0 + chunk);
    int cnt = 0;
    bool bad = false;
    for (int i = r0; i < r1; ++i) {
        const int t = residue_type(seq[i]);
        if (t < 0) {
            bad = true;
        } else {
            if (s_ra) s_ra[i] = cnt | (t << 24);
            cnt += res_i[t] & 0xff;
        }
    }
    if (bad) atomicOr(&s_flag[1], (int)TPL_A2C_STATUS_SEQ);
    // atoms of this run and the later ones: prefix over threads, by DPP
    // within the wave and the earlier waves' totals through LDS
    constexpr int NW = NT / 64;
    const int wv = tid >> 6;
    const int cinc = wave_scan_int(cnt);
    if ((tid & 63) == 63) s_c[wv] = cinc;
    __syncthreads();
    int before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        if (w < wv) before += s_c[w];
        total += s_c[w];
    }
    run.total = total;
    run.atom0 = total - (before + cinc);  // atoms of the runs before this one
    if (s_ra)
        for (int i = r0; i < r1; ++i) s_ra[i] += run.atom0;
    if (run.total > atoms_stride && tid == 0) atomicOr(&s_flag[1], (int)TPL_A2C_STATUS_ATOMS);
    __syncthreads();
    const int flags = s_flag[1];
    if (flags && tid == 0 && status)
        __hip_atomic_fetch_or(status, (uint32_t)flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    run.L = L;
    run.r0 = r0;
    run.r1 = r1;
    if (flags) run.L = run.r0 = run.r1 = run.total = 0;
}

template <typename T>
__device__ __forceinline__ void load3(const T* p, double v[3]) {
    v[0] = (double)p[0];
    v[1] = (double)p[1];
    v[2] = (double)p[2];
}

// Sums over atoms [k0, k1) of a structure.  With UNROLL, four atoms at a time
// with their loads issued first.  This kernel now serves small batches,
// where it is latency-bound: at 512 threads, 32 x 350 residues 18.0 -> 16.9
// us, one structure of 1000 36.0 -> 33.9, 64 x 1000 43.9 -> 41.0.  At 256
// threads the extra registers cost a wave per SIMD (1024 x 30: 13.9 -> 18.7
// us), so it stays off there.  (When this kernel also served B=1024 of
// mixed lengths 30-1000, unrolling measured slower there: 153 -> 204 us.)
template <bool UNROLL, typename T>
__device__ __forceinline__ void atom_sums(const T* xs, const T* gs, int k0, int k1, Sums& acc) {
    int k = k0;
    if (UNROLL) {
        for (; k + 4 <= k1; k += 4) {
            double x[4][3], g[4][3];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                load3(xs + 3 * (size_t)(k + j), x[j]);
                load3(gs + 3 * (size_t)(k + j), g[j]);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) sums_add(acc, x[j], g[j]);
        }
    }
    for (; k < k1; ++k) {
        double x[3], g[3];
        load3(xs + 3 * (size_t)k, x);
        load3(gs + 3 * (size_t)k, g);
        sums_add(acc, x, g);
    }
}

// u . (xg - p cross g) with u the unit vector from q to p.
__device__ __forceinline__ double bond_torque(const double p[3], const double q[3], const Sums& s) {
    const double d0 = p[0] - q[0], d1 = p[1] - q[1], d2 = p[2] - q[2];
    const double inv = 1.0 / sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    const double v0 = s.xg[0] - (p[1] * s.g[2] - p[2] * s.g[1]);
    const double v1 = s.xg[1] - (p[2] * s.g[0] - p[0] * s.g[2]);
    const double v2 = s.xg[2] - (p[0] * s.g[1] - p[1] * s.g[0]);
    return (d0 * v0 + d1 * v1 + d2 * v2) * inv;
}

template <int NT, typename T>
__global__ void __launch_bounds__(NT) a2c_bwd(const uint8_t* __restrict__ seqs, int seq_stride,
                                              const T* __restrict__ coords, const T* __restrict__ grad_coords,
                                              int atoms_stride, T* __restrict__ grad_angles, int ang_stride,
                                              uint32_t* status, const int* __restrict__ order) {
    __shared__ int s_res[kTypes];
    __shared__ int s_side[kTypes * kMaxSide];
    __shared__ double s_v[6][NT / 64];  // the waves' sums
    __shared__ int s_c[NT / 64];        // the waves' atom counts
    __shared__ int s_flag[2];
    const int b = structure_of(order, status), tid = threadIdx.x;
    if (b < 0) return;
    const uint8_t* seq = seqs + (size_t)b * seq_stride;
    const T* xs = coords + (size_t)b * 3 * atoms_stride;
    const T* gs = grad_coords + (size_t)b * 3 * atoms_stride;
    T* ga = grad_angles + (size_t)b * 8 * ang_stride;
    for (int k = tid; k < kTypes * kMaxSide; k += NT) {
        const Side sd = kSide[k / kMaxSide][k % kMaxSide];
        s_side[k] = (sd.row & 0xff) | (sd.first << 8) | (sd.count << 16);
    }
    if (tid < kTypes) {
        const Residue r = kRes[tid];
        s_res[tid] = r.natoms | (r.nca << 8) | (r.nside << 16);
    }
    __syncthreads();
    Run run;
    count_run<NT>(seq, seq_stride, ang_stride, atoms_stride, status, s_res, s_c, s_flag, run);
    // pass A: this run's (sum g, sum x cross g)
    Sums run_sum;
    sums_zero(run_sum);
    {
        int a = run.atom0;
        for (int i = run.r0; i < run.r1; ++i) a += s_res[residue_type(seq[i])] & 0xff;
        atom_sums<NT == kThreadsLarge>(xs, gs, run.atom0, a, run_sum);
    }
    // incl = sums over every atom from this run on: runs are in reverse
    // thread order, so an inclusive prefix sum over threads (DPP within the
    // wave, then the earlier waves' totals through LDS)
    double v[6];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        v[c] = run_sum.g[c];
        v[3 + c] = run_sum.xg[c];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = wave_scan_double(v[k]);
    if ((tid & 63) == 63) {
#pragma unroll
        for (int k = 0; k < 6; ++k) s_v[k][tid >> 6] = v[k];
    }
    __syncthreads();
    for (int w = 0; w < (tid >> 6); ++w) {
#pragma unroll
        for (int k = 0; k < 6; ++k) v[k] += s_v[k][w];
    }
    Sums incl;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        incl.g[c] = v[c];
        incl.xg[c] = v[3 + c];
    }
    // pass B: gradients, residue by residue
    int a = run.atom0;
    for (int i = run.r0; i < run.r1; ++i) {
        const int t = residue_type(seq[i]);
        const int ri = s_res[t];
        const int natoms = ri & 0xff, nca = (ri >> 8) & 0xff, nside = ri >> 16;
        Sums sN, sCA, sSide, sC, excl, s;
        sums_zero(sN);
        sums_zero(sCA);
        sums_zero(sSide);
        sums_zero(sC);
        atom_sums<NT == kThreadsLarge>(xs, gs, a, a + 1, sN);
        atom_sums<NT == kThreadsLarge>(xs, gs, a + 1, a + 1 + nca, sCA);
        atom_sums<NT == kThreadsLarge>(xs, gs, a + natoms - 2, a + natoms, sC);
        double pN[3], pCA[3], pC[3];
        load3(xs + 3 * (size_t)a, pN);
        load3(xs + 3 * (size_t)(a + 1), pCA);
        load3(xs + 3 * (size_t)(a + natoms - 2), pC);
        for (int r = 3; r < 8; ++r) ga[(size_t)r * ang_stride + i] = (T)0;
        // side chain, last node first, so that sSide holds node n's group and
        // the later nodes' groups when node n is reached (each atom read
        // once).  Node n hangs on node n-1, or on CA when n-1 is the fixed
        // first node (no atoms, at CA's origin; every other node has atoms).
        for (int n = nside - 1; n >= 0; --n) {
            const int si = s_side[t * kMaxSide + n];
            const int row = (int)(int8_t)(si & 0xff), first = (si >> 8) & 0xff, count = si >> 16;
            if (count == 0) continue;
            atom_sums<NT == kThreadsLarge>(xs, gs, a + first, a + first + count, sSide);
            if (row >= 0) {
                const int sp = n > 0 ? s_side[t * kMaxSide + n - 1] : 0;
                const int pfirst = (sp >> 8) & 0xff, pcount = sp >> 16;
                double p[3], q[3];
                load3(xs + 3 * (size_t)(a + first), p);
                if (pcount > 0)
                    load3(xs + 3 * (size_t)(a + pfirst), q);
                else
                    q[0] = pCA[0], q[1] = pCA[1], q[2] = pCA[2];
                ga[(size_t)row * ang_stride + i] = (T)bond_torque(p, q, sSide);
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            excl.g[c] = incl.g[c] - (sN.g[c] + sCA.g[c] + sSide.g[c] + sC.g[c]);
            excl.xg[c] = incl.xg[c] - (sN.xg[c] + sCA.xg[c] + sSide.xg[c] + sC.xg[c]);
        }
        // omega of residue i-1 binds N_i (on the previous C): the whole residue and everything after
        if (i > 0) {
            double pCp[3];
            load3(xs + 3 * (size_t)(a - 2), pCp);
            ga[(size_t)2 * ang_stride + i - 1] = (T)bond_torque(pN, pCp, incl);
        }
        if (i == run.L - 1) ga[(size_t)2 * ang_stride + i] = (T)0;
        // phi binds CA (on N): all but N
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            s.g[c] = incl.g[c] - sN.g[c];
            s.xg[c] = incl.xg[c] - sN.xg[c];
        }
        ga[i] = (T)bond_torque(pCA, pN, s);
        // psi binds C (on CA): the C group and everything after
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            s.g[c] = sC.g[c] + excl.g[c];
            s.xg[c] = sC.xg[c] + excl.xg[c];
        }
        ga[(size_t)ang_stride + i] = (T)bond_torque(pC, pCA, s);
        incl = excl;
        a += natoms;
    }
    // angles past the sequence (and every angle of a flagged structure)
    for (int k = tid; k < 8 * ang_stride; k += NT)
        if (k % ang_stride >= run.L) ga[k] = (T)0;
}

// ---- backward over atom tiles ----
// The run-per-thread backward above makes every load of a wave touch 64
// cache lines (the lanes' runs lie far apart), and it reads each atom twice.
// This one walks a structure from its end in tiles of up to kTileAtoms
// consecutive atoms, cut at residue starts: thread t takes the tile's atoms
// hi-1-Kt ... hi-K-Kt, so a wave's loads cover consecutive atoms, and each
// atom is read once.  Per tile: a workgroup prefix scan over threads (in
// descending atom order, so a suffix over atoms) of (sum g, sum x cross g)
// plus the carry from the later tiles gives the suffix sum SUF[k] at every
// atom; every node's subtree is a suffix (backbone nodes: from the node's
// origin on) or a suffix minus the suffix at the residue's C (side-chain
// nodes: side groups are contiguous and end right before C), so the thread
// holding a node's origin atom writes that node's gradient.  The next tile's
// atoms are loaded while the current one is processed.
constexpr int kTileThreads = 512;
constexpr int kTileK = 2;                                   // atoms per thread and tile
constexpr int kTileAtoms = kTileThreads * kTileK;           // 1024
constexpr int kTileRes = kTileAtoms / 4;                    // residues per tile at most (4 atoms or more each)
constexpr int kTileStride = kTileAtoms - (kMaxAtoms - 1);   // tile p starts at the first residue starting at or after p * this
constexpr int kTileMaxRes = 2048;                           // residues per structure (the residue starts live in LDS)
constexpr int kTileMaxTiles = kTileMaxRes * kMaxAtoms / kTileStride + 1;

// 1 / sqrt(x) for x > 0: the hardware estimate and two Newton steps
// (each squares the relative error).
__device__ __forceinline__ double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    y = y * (1.5 - h * y * y);
    y = y * (1.5 - h * y * y);
    return y;
}

// u . (xg - p cross g) with u the unit vector from q to p (bond_torque with
// the reciprocal square root above).
__device__ __forceinline__ double bond_torque_rsq(const double p[3], const double q[3], const double s[6]) {
    const double d0 = p[0] - q[0], d1 = p[1] - q[1], d2 = p[2] - q[2];
    const double v0 = s[3] - (p[1] * s[2] - p[2] * s[1]);
    const double v1 = s[4] - (p[2] * s[0] - p[0] * s[2]);
    const double v2 = s[5] - (p[0] * s[1] - p[1] * s[0]);
    return (d0 * v0 + d1 * v1 + d2 * v2) * rsqrt_nr(d0 * d0 + d1 * d1 + d2 * d2);
}

// The tile's atoms of this thread; atoms below lo read atom lo (in bounds)
// and are zeroed.
template <typename T>
__device__ __forceinline__ void load_tile_atoms(const T* xs, const T* gs, int lo, int hi, T (&x)[kTileK][3],
                                                T (&g)[kTileK][3]) {
#pragma unroll
    for (int m = 0; m < kTileK; ++m) {
        const int k = hi - 1 - (kTileK * (int)threadIdx.x + m);
        const size_t kc = 3 * (size_t)(k >= lo ? k : lo);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            x[m][c] = xs[kc + c];
            g[m][c] = gs[kc + c];
        }
        if (k < lo) {
#pragma unroll
            for (int c = 0; c < 3; ++c) x[m][c] = g[m][c] = (T)0;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(4, 8))) a2c_bwd_tiled(const uint8_t* __restrict__ seqs, int seq_stride,
                                                              const T* __restrict__ coords,
                                                              const T* __restrict__ grad_coords, int atoms_stride,
                                                              T* __restrict__ grad_angles, int ang_stride,
                                                              uint32_t* status, const int* __restrict__ order) {
    constexpr int NT = kTileThreads, NW = NT / 64, K = kTileK;
    __shared__ int s_res[kTypes];                   // natoms | nca << 8 | nside << 16 | bound rows << 24
    __shared__ int s_side[kTypes * kMaxSide];       // row & 0xff | first << 8 | count << 16
    __shared__ uint32_t s_slot[kTypes * kMaxAtoms]; // per (type, slot): the role bits (tile_roles)
    __shared__ int s_ra[kTileMaxRes];               // residue's first atom | type << 24
    __shared__ int s_tile[kTileMaxTiles];           // each tile's first residue
    __shared__ uint32_t s_role[2][kTileAtoms];      // tile atoms' roles (tile_roles)
    __shared__ T s_x[2][3 * kTileAtoms];            // tile atoms' positions
    __shared__ double s_cs[kTileRes][6];            // suffix sums at each tile residue's C
    __shared__ double s_v[6][NW];                   // the waves' sums
    __shared__ int s_c[NW];
    __shared__ int s_flag[2];
    const int b = structure_of(order, status), tid = threadIdx.x;
    if (b < 0) return;
    const uint8_t* seq = seqs + (size_t)b * seq_stride;
    const T* xs = coords + (size_t)b * 3 * atoms_stride;
    const T* gs = grad_coords + (size_t)b * 3 * atoms_stride;
    T* ga = grad_angles + (size_t)b * 8 * ang_stride;
    for (int k = tid; k < kTypes * kMaxSide; k += NT) {
        const Side sd = kSide[k / kMaxSide][k % kMaxSide];
        s_side[k] = (sd.row & 0xff) | (sd.first << 8) | (sd.count << 16);
    }
    if (tid < kTypes) {
        const Residue r = kRes[tid];
        s_res[tid] = r.natoms | (r.nca << 8) | (r.nside << 16);
    }
    __syncthreads();
    // Per (type t, atom slot s), the role bits of tile_roles but the
    // residue's index: whether the atom is the origin of a node with an
    // angle, the angle's row, whether that angle belongs to the previous
    // residue (omega, on N, from the previous C two atoms back), how many
    // atoms back the parent node's origin lies, whether the node is a side
    // one, whether the atom is C.  A side node hangs on the node before it,
    // or on CA when that one is the fixed first node without atoms.  And per
    // type, the side rows (3..7) it writes: nodes with atoms and a bound
    // angle.
    for (int k = tid; k < kTypes * kMaxAtoms; k += NT) {
        const int t = k / kMaxAtoms, s = k % kMaxAtoms;
        const int natoms = s_res[t] & 0xff, nside = (s_res[t] >> 16) & 0xff;
        uint32_t role = s == natoms - 2 ? 1u << 18 : 0u;
        if (s == 0)
            role |= (1u << 8) | (2u << 9) | (1u << 12) | (2u << 13);  // omega: N on the previous C
        else if (s == 1)
            role |= (1u << 8) | (0u << 9) | (1u << 13);  // phi: CA on N
        else if (s == natoms - 2)
            role |= (1u << 8) | (1u << 9) | ((uint32_t)(s - 1) << 13);  // psi: C on CA
        else if (s < natoms)
            for (int n = 0; n < nside; ++n) {
                const int si = s_side[t * kMaxSide + n];
                const int row = (int)(int8_t)(si & 0xff), first = (si >> 8) & 0xff, count = si >> 16;
                if (count > 0 && row >= 0 && first == s) {
                    const int sp = n > 0 ? s_side[t * kMaxSide + n - 1] : 0;
                    const int parent = (sp >> 16) > 0 ? (sp >> 8) & 0xff : 1;
                    role |= (1u << 8) | ((uint32_t)row << 9) | ((uint32_t)(s - parent) << 13) | (1u << 17);
                }
            }
        s_slot[k] = role;
    }
    int bound = 0;
    if (tid < kTypes) {
        for (int n = 0; n < ((s_res[tid] >> 16) & 0xff); ++n) {
            const int si = s_side[tid * kMaxSide + n];
            if ((si >> 16) > 0 && (int)(int8_t)(si & 0xff) >= 0) bound |= 1 << (si & 0xff);
        }
    }
    __syncthreads();
    if (tid < kTypes) s_res[tid] |= bound << 24;
    Run run;
    count_run<NT>(seq, seq_stride, ang_stride, atoms_stride, status, s_res, s_c, s_flag, run, s_ra);
    __syncthreads();
    const int L = run.L;
    for (int i = tid; i < L; i += NT) {
        const int p = (s_ra[i] & 0xffffff) / kTileStride;
        if (i == 0 || p != (s_ra[i - 1] & 0xffffff) / kTileStride) s_tile[p] = i;
    }
    __syncthreads();
    const int ntiles = L > 0 ? (s_ra[L - 1] & 0xffffff) / kTileStride + 1 : 0;
    const size_t stride = (size_t)ang_stride;
    // Roles of tile p's atoms into s_role[p & 1], one residue per thread:
    // bits 0-7 the residue's index in the tile, bit 8 set if the atom is the
    // origin of a node with an angle, bits 9-11 that angle's row, bit 12 if
    // it belongs to the previous residue (omega, on N), bits 13-16 how many
    // atoms back the parent node's origin lies, bit 17 for side-chain nodes
    // (the suffix at C is subtracted), bit 18 if the atom is the residue's
    // C.  The thread also writes the residue's angles that no node carries:
    // the side rows its type does not bind, and omega of the last residue.
    auto tile_roles = [&](int p) {
        const int ilo = s_tile[p], ihi = p + 1 < ntiles ? s_tile[p + 1] : L;
        const int lo = s_ra[ilo] & 0xffffff;
        if (tid < ihi - ilo) {
            const int i = ilo + tid, ra = s_ra[i];
            const int a = ra & 0xffffff, t = ra >> 24, rt = s_res[t], n = rt & 0xff;
#pragma unroll
            for (int s = 0; s < kMaxAtoms; ++s)
                if (s < n) {
                    uint32_t role = s_slot[t * kMaxAtoms + s] | (uint32_t)tid;
                    if (s == 0 && i == 0) role &= ~(1u << 8);  // the first residue has no omega before it
                    s_role[p & 1][a - lo + s] = role;
                }
            const int bnd = (int)((uint32_t)rt >> 24);
            for (int r = 3; r < 8; ++r)
                if (!((bnd >> r) & 1)) ga[r * stride + i] = (T)0;
            if (i == L - 1) ga[2 * stride + i] = (T)0;
        }
    };
    double carry[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) carry[c] = 0.0;
    T cx[K][3], cg[K][3];
    int lo = 0, hi = 0;
    if (ntiles > 0) {
        lo = s_ra[s_tile[ntiles - 1]] & 0xffffff;
        hi = run.total;
        load_tile_atoms(xs, gs, lo, hi, cx, cg);
        tile_roles(ntiles - 1);
        __syncthreads();
    }
    for (int p = ntiles - 1; p >= 0; --p) {
        const int buf = p & 1;
        const int ilo = s_tile[p];
        // the previous residue's C (omega of residue ilo - 1 turns about the
        // bond from it), then the next (lower) tile's atoms
        T qprev[3] = {(T)0, (T)0, (T)0};
        if (ilo > 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) qprev[c] = xs[3 * (size_t)(lo - 2) + c];
        }
        T nx[K][3], ng[K][3];
        const int nlo = p > 0 ? s_ra[s_tile[p - 1]] & 0xffffff : 0;
        load_tile_atoms(xs, gs, nlo, p > 0 ? lo : nlo, nx, ng);
        // this tile's positions, for the parent origins; the next tile's roles
#pragma unroll
        for (int m = 0; m < K; ++m) {
            const int k = hi - 1 - (K * tid + m);
            if (k >= lo) {
#pragma unroll
                for (int c = 0; c < 3; ++c) s_x[buf][3 * (k - lo) + c] = cx[m][c];
            }
        }
        if (p > 0) tile_roles(p - 1);
        // per atom: the node whose origin it is, its angle's slot in ga, its
        // parent's origin in s_x (-1: qprev), whether the suffix at C is
        // subtracted, and whether the atom is a C
        int dst[K], qi[K], il[K];
        bool sub[K], isc[K];
#pragma unroll
        for (int m = 0; m < K; ++m) {
            const int k = hi - 1 - (K * tid + m);
            const uint32_t role = k >= lo ? s_role[buf][k - lo] : 0u;
            il[m] = role & 0xff;
            dst[m] = (role >> 8) & 1 ? (int)((role >> 9) & 7) * ang_stride + ilo + il[m] - (int)((role >> 12) & 1) : -1;
            qi[m] = k - (int)((role >> 13) & 15) - lo;
            sub[m] = (role >> 17) & 1;
            isc[m] = k >= lo && ((role >> 18) & 1);
        }
        // (g, x cross g) per atom; the thread's sum; prefix over threads
        double w[K][6], tot[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) tot[c] = 0.0;
#pragma unroll
        for (int m = 0; m < K; ++m) {
            const double x0 = (double)cx[m][0], x1 = (double)cx[m][1], x2 = (double)cx[m][2];
            const double g0 = (double)cg[m][0], g1 = (double)cg[m][1], g2 = (double)cg[m][2];
            w[m][0] = g0;
            w[m][1] = g1;
            w[m][2] = g2;
            w[m][3] = x1 * g2 - x2 * g1;
            w[m][4] = x2 * g0 - x0 * g2;
            w[m][5] = x0 * g1 - x1 * g0;
#pragma unroll
            for (int c = 0; c < 6; ++c) tot[c] += w[m][c];
        }
        double incl[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) incl[c] = wave_scan_double(tot[c]);
        if ((tid & 63) == 63) {
#pragma unroll
            for (int c = 0; c < 6; ++c) s_v[c][tid >> 6] = incl[c];
        }
        __syncthreads();
        // SUF at the thread's atoms: carry (later tiles) + later waves +
        // later lanes + the thread's atoms from the highest down.  Lane v of
        // every wave takes wave v's total and a row scan makes them
        // inclusive over waves (NW <= 16), read back with readlane.
        const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
        double wt[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) wt[c] = row_scan_double(lane < NW ? s_v[c][lane] : 0.0);
        double suf[K][6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const double all = readlane_double(wt[c], NW - 1);
            const double before = wv > 0 ? readlane_double(wt[c], wv - 1) : 0.0;
            double acc = carry[c] + before + (incl[c] - tot[c]);
#pragma unroll
            for (int m = 0; m < K; ++m) {
                acc += w[m][c];
                suf[m][c] = acc;
            }
            carry[c] += all;
        }
#pragma unroll
        for (int m = 0; m < K; ++m)
            if (isc[m]) {
#pragma unroll
                for (int c = 0; c < 6; ++c) s_cs[il[m]][c] = suf[m][c];
            }
        __syncthreads();
        // gradients of the nodes whose origins this thread holds
#pragma unroll
        for (int m = 0; m < K; ++m) {
            if (dst[m] < 0) continue;
            const double p3[3] = {(double)cx[m][0], (double)cx[m][1], (double)cx[m][2]};
            double q[3], sm[6];
            const int qk = qi[m] >= 0 ? qi[m] : 0;
#pragma unroll
            for (int c = 0; c < 3; ++c) q[c] = (double)(qi[m] >= 0 ? s_x[buf][3 * qk + c] : qprev[c]);
#pragma unroll
            for (int c = 0; c < 6; ++c) sm[c] = sub[m] ? suf[m][c] - s_cs[il[m]][c] : suf[m][c];
            ga[dst[m]] = (T)bond_torque_rsq(p3, q, sm);
        }
        hi = lo;
        lo = nlo;
#pragma unroll
        for (int m = 0; m < K; ++m)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                cx[m][c] = nx[m][c];
                cg[m][c] = ng[m][c];
            }
    }
    // angles past the sequence (and every angle of a flagged structure)
    for (int k = tid; k < 8 * ang_stride; k += NT)
        if (k % ang_stride >= L) ga[k] = (T)0;
}

// 512-thread workgroups when some structure may be longer than 256 residues:
// always for the run backward (mixed lengths 30-1000 at B=1024: 155 -> 129
// us), and for the forward only up to twice as many structures as CUs (B=256
// of 1000 residues: 33.7 -> 28.6 us; B=512 of 300: 23.8 -> 21.8 us; at
// B=1024 512 threads measured slower, with fewer workgroups per CU: mixed
// lengths 30-1000 48.4 -> 55.7 us, 1024 x 270 32.0 -> 39.6 us).
bool use_large_fwd(int batch, int ang_stride) {
    return batch <= 512 && ang_stride > kThreadsSmall;
}

bool use_large_bwd(int ang_stride) {
    return ang_stride > kThreadsSmall;
}

// The atom-tile backward where the batch fills the GPU: it moves the atoms
// with coalesced loads and reads each once, but walks a structure's tiles
// one after another, so a lone long structure takes longer than with
// per-thread runs.  Measured (FP32, backward us, runs -> tiles): mixed
// lengths 30-1000 at B=1024 108 -> 87, 256 x 1000 88 -> 49, 128 x 1000
// 52 -> 48, 512 x 300 34 -> 25, 1024 x 270 56 -> 47; 64 x 1000 44 -> 47,
// 128 x 300 17.5 -> 18.9, 32 x 350 17.8 -> 18.9, 1024 x 30 13.7 -> 21.4,
// 1 x 1000 36 -> 45.
bool use_tiles(int batch, int ang_stride) {
    return ang_stride <= kTileMaxRes && (int64_t)batch * ang_stride >= 100000;
}

// g_tab of the current device, filled on the first forward there: the
// kernel runs on the caller's stream, which is then synchronised once, so
// that forwards on any stream of the device find the tables.
bool tables_ready(hipStream_t s) {
    static std::mutex mu;
    static std::vector<char> ready;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return false;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)ready.size() <= dev) ready.resize(dev + 1, 0);
    if (ready[dev]) return true;
    hipLaunchKernelGGL(a2c_init_tables, dim3(1), dim3(256), 0, s);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) return false;
    ready[dev] = 1;
    return true;
}

int check(int dtype, int B, int seq_stride, int ang_stride, int atoms_stride) {
    if (dtype != TPL_A2C_F32 && dtype != TPL_A2C_F64) return TPL_A2C_ERR_ARG;
    if (B < 0 || seq_stride < 0 || ang_stride < 0 || atoms_stride < 0) return TPL_A2C_ERR_ARG;
    if ((int64_t)8 * ang_stride > INT32_MAX || (int64_t)3 * atoms_stride > INT32_MAX)
        return TPL_A2C_ERR_SIZE;
    return TPL_A2C_OK;
}

}  // namespace

extern "C" {

const char* tpl_a2c_strerror(int code) {
    switch (code) {
        case TPL_A2C_OK: return "ok";
        case TPL_A2C_ERR_ARG: return "invalid argument (null pointer, negative size or unknown dtype)";
        case TPL_A2C_ERR_LAUNCH: return "kernel launch failed";
        case TPL_A2C_ERR_SIZE: return "sizes too large";
        default: return "unknown error code";
    }
}

int tpl_a2c_num_atoms(const char* sequence) {
    static const Residue res[kTypes] = A2C_RESIDUES;
    if (!sequence) return -1;
    int n = 0;
    for (const char* p = sequence; *p; ++p) {
        const int t = residue_type((uint8_t)*p);
        if (t < 0) return -1;
        n += res[t].natoms;
    }
    return n;
}

int tpl_a2c_forward_ordered(const int32_t* order, int dtype, const uint8_t* sequences, int seq_stride,
                            const void* angles, int angles_stride, void* coords, uint8_t* res_names,
                            uint8_t* atom_names, int atoms_stride, uint32_t* status, int batch,
                            tpl_a2c_stream_t stream) {
    const int rc = check(dtype, batch, seq_stride, angles_stride, atoms_stride);
    if (rc != TPL_A2C_OK) return rc;
    if (batch == 0) return TPL_A2C_OK;
    if (!sequences || (angles_stride > 0 && !angles) || (atoms_stride > 0 && !coords) ||
        (!res_names) != (!atom_names))
        return TPL_A2C_ERR_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!tables_ready(s)) return TPL_A2C_ERR_LAUNCH;
    uint32_t* rn = reinterpret_cast<uint32_t*>(res_names);
    uint32_t* an = reinterpret_cast<uint32_t*>(atom_names);
    const bool large = use_large_fwd(batch, angles_stride);
    if (dtype == TPL_A2C_F32) {
        const float* a = static_cast<const float*>(angles);
        float* c = static_cast<float*>(coords);
        if (large)
            hipLaunchKernelGGL((a2c_fwd<kThreadsLarge, float>), dim3(batch), dim3(kThreadsLarge), 0, s,
                               sequences, seq_stride, a, angles_stride, c, rn, an, atoms_stride, status, order);
        else
            hipLaunchKernelGGL((a2c_fwd<kThreadsSmall, float>), dim3(batch), dim3(kThreadsSmall), 0, s,
                               sequences, seq_stride, a, angles_stride, c, rn, an, atoms_stride, status, order);
    } else {
        const double* a = static_cast<const double*>(angles);
        double* c = static_cast<double*>(coords);
        if (large)
            hipLaunchKernelGGL((a2c_fwd<kThreadsLarge, double>), dim3(batch), dim3(kThreadsLarge), 0, s,
                               sequences, seq_stride, a, angles_stride, c, rn, an, atoms_stride, status, order);
        else
            hipLaunchKernelGGL((a2c_fwd<kThreadsSmall, double>), dim3(batch), dim3(kThreadsSmall), 0, s,
                               sequences, seq_stride, a, angles_stride, c, rn, an, atoms_stride, status, order);
    }
    return hipGetLastError() == hipSuccess ? TPL_A2C_OK : TPL_A2C_ERR_LAUNCH;
}

int tpl_a2c_forward(int dtype, const uint8_t* sequences, int seq_stride, const void* angles,
                    int angles_stride, void* coords, uint8_t* res_names, uint8_t* atom_names,
                    int atoms_stride, uint32_t* status, int batch, tpl_a2c_stream_t stream) {
    return tpl_a2c_forward_ordered(nullptr, dtype, sequences, seq_stride, angles, angles_stride, coords,
                                   res_names, atom_names, atoms_stride, status, batch, stream);
}

int tpl_a2c_backward_form(int form, const int32_t* order, int dtype, const uint8_t* sequences, int seq_stride,
                          const void* coords, const void* grad_coords, int atoms_stride, void* grad_angles,
                          int angles_stride, uint32_t* status, int batch, tpl_a2c_stream_t stream) {
    if (form != TPL_A2C_BWD_AUTO && form != TPL_A2C_BWD_RUNS && form != TPL_A2C_BWD_TILES) return TPL_A2C_ERR_ARG;
    const int rc = check(dtype, batch, seq_stride, angles_stride, atoms_stride);
    if (rc != TPL_A2C_OK) return rc;
    if (form == TPL_A2C_BWD_TILES && angles_stride > kTileMaxRes) return TPL_A2C_ERR_SIZE;
    if (batch == 0 || angles_stride == 0) return TPL_A2C_OK;
    if (!sequences || !grad_angles || (atoms_stride > 0 && (!coords || !grad_coords)))
        return TPL_A2C_ERR_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (form == TPL_A2C_BWD_AUTO) form = use_tiles(batch, angles_stride) ? TPL_A2C_BWD_TILES : TPL_A2C_BWD_RUNS;
    if (form == TPL_A2C_BWD_TILES) {
        if (dtype == TPL_A2C_F32)
            hipLaunchKernelGGL((a2c_bwd_tiled<float>), dim3(batch), dim3(kTileThreads), 0, s, sequences,
                               seq_stride, static_cast<const float*>(coords),
                               static_cast<const float*>(grad_coords), atoms_stride,
                               static_cast<float*>(grad_angles), angles_stride, status, order);
        else
            hipLaunchKernelGGL((a2c_bwd_tiled<double>), dim3(batch), dim3(kTileThreads), 0, s, sequences,
                               seq_stride, static_cast<const double*>(coords),
                               static_cast<const double*>(grad_coords), atoms_stride,
                               static_cast<double*>(grad_angles), angles_stride, status, order);
        return hipGetLastError() == hipSuccess ? TPL_A2C_OK : TPL_A2C_ERR_LAUNCH;
    }
    const bool large = use_large_bwd(angles_stride);
    if (dtype == TPL_A2C_F32) {
        const float *x = static_cast<const float*>(coords), *g = static_cast<const float*>(grad_coords);
        float* ga = static_cast<float*>(grad_angles);
        if (large)
            hipLaunchKernelGGL((a2c_bwd<kThreadsLarge, float>), dim3(batch), dim3(kThreadsLarge), 0, s,
                               sequences, seq_stride, x, g, atoms_stride, ga, angles_stride, status, order);
        else
            hipLaunchKernelGGL((a2c_bwd<kThreadsSmall, float>), dim3(batch), dim3(kThreadsSmall), 0, s,
                               sequences, seq_stride, x, g, atoms_stride, ga, angles_stride, status, order);
    } else {
        const double *x = static_cast<const double*>(coords), *g = static_cast<const double*>(grad_coords);
        double* ga = static_cast<double*>(grad_angles);
        if (large)
            hipLaunchKernelGGL((a2c_bwd<kThreadsLarge, double>), dim3(batch), dim3(kThreadsLarge), 0, s,
                               sequences, seq_stride, x, g, atoms_stride, ga, angles_stride, status, order);
        else
            hipLaunchKernelGGL((a2c_bwd<kThreadsSmall, double>), dim3(batch), dim3(kThreadsSmall), 0, s,
                               sequences, seq_stride, x, g, atoms_stride, ga, angles_stride, status, order);
    }
    return hipGetLastError() == hipSuccess ? TPL_A2C_OK : TPL_A2C_ERR_LAUNCH;
}

int tpl_a2c_backward(int dtype, const uint8_t* sequences, int seq_stride, const void* coords,
                     const void* grad_coords, int atoms_stride, void* grad_angles, int angles_stride,
                     uint32_t* status, int batch, tpl_a2c_stream_t stream) {
    return tpl_a2c_backward_form(TPL_A2C_BWD_AUTO, nullptr, dtype, sequences, seq_stride, coords, grad_coords,
                                 atoms_stride, grad_angles, angles_stride, status, batch, stream);
}

}  // extern "C"
