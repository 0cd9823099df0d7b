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
// part, and each thread walks its residues once more.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "a2c_table.h"
#include "sincos_tab.h"
#include "tpl_a2c.h"

namespace {

using namespace a2c;

// Workgroup sizes: 256 threads, or 512 for long structures (see use_large_fwd
// and use_large_bwd), where a thread then walks half as many residues.  LDS:
// forward 13.6 / 14.0 KB (the tables), backward 0.8 / 1.0 KB.
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
                                                    uint32_t* status) {
    __shared__ Tables tb;
    __shared__ double s_m[12][NT / 64];  // the waves' products
    __shared__ int s_c[NT / 64];          // the waves' atom counts
    __shared__ int s_flag[2];
    const int b = blockIdx.x;
    const uint8_t* seq = seqs + (size_t)b * seq_stride;
    const T* ang = angles + (size_t)b * 8 * ang_stride;
    T* out = coords + (size_t)b * 3 * atoms_stride;
    fill_tables<NT>(tb);
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
template <int NT>
__device__ void count_run(const uint8_t* seq, int seq_stride, int ang_stride, int atoms_stride,
                          uint32_t* status, const int* res_i, int* s_c, int* s_flag, Run& run) {
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
        if (t < 0)
            bad = true;
        else
            cnt += res_i[t] & 0xff;
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

// Sums over atoms [k0, k1) of a structure.  (Unrolling by four atoms with
// the loads issued first measured slower: 153 -> 204 us for the mixed1000
// backward with it in pass A only, 281 us with it everywhere.)
template <typename T>
__device__ __forceinline__ void atom_sums(const T* xs, const T* gs, int k0, int k1, Sums& acc) {
    for (int k = k0; k < k1; ++k) {
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
                                              uint32_t* status) {
    __shared__ int s_res[kTypes];
    __shared__ int s_side[kTypes * kMaxSide];
    __shared__ double s_v[6][NT / 64];  // the waves' sums
    __shared__ int s_c[NT / 64];        // the waves' atom counts
    __shared__ int s_flag[2];
    const int b = blockIdx.x, tid = threadIdx.x;
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
        atom_sums(xs, gs, run.atom0, a, run_sum);
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
        atom_sums(xs, gs, a, a + 1, sN);
        atom_sums(xs, gs, a + 1, a + 1 + nca, sCA);
        atom_sums(xs, gs, a + natoms - 2, a + natoms, sC);
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
            atom_sums(xs, gs, a + first, a + first + count, sSide);
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

// 512-thread workgroups when some structure may be longer than 256 residues:
// always for the backward (mixed lengths 30-1000 at B=1024: 155 -> 129 us),
// and for the forward only when the structures are at most as many as the
// CUs (B=256 of 1000 residues: 33.7 -> 28.6 us; at B=1024 with mixed lengths
// 512 threads measured slower, 57 -> 72 us, with fewer workgroups per CU).
bool use_large_fwd(int batch, int ang_stride) {
    return batch <= 256 && ang_stride > kThreadsSmall;
}

bool use_large_bwd(int ang_stride) {
    return ang_stride > kThreadsSmall;
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

int tpl_a2c_forward(int dtype, const uint8_t* sequences, int seq_stride, const void* angles,
                    int angles_stride, void* coords, uint8_t* res_names, uint8_t* atom_names,
                    int atoms_stride, uint32_t* status, int batch, tpl_a2c_stream_t stream) {
    const int rc = check(dtype, batch, seq_stride, angles_stride, atoms_stride);
    if (rc != TPL_A2C_OK) return rc;
    if (batch == 0) return TPL_A2C_OK;
    if (!sequences || (angles_stride > 0 && !angles) || (atoms_stride > 0 && !coords) ||
        (!res_names) != (!atom_names))
        return TPL_A2C_ERR_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint32_t* rn = reinterpret_cast<uint32_t*>(res_names);
    uint32_t* an = reinterpret_cast<uint32_t*>(atom_names);
    const bool large = use_large_fwd(batch, angles_stride);
    if (dtype == TPL_A2C_F32) {
        const float* a = static_cast<const float*>(angles);
        float* c = static_cast<float*>(coords);
        if (large)
            hipLaunchKernelGGL((a2c_fwd<kThreadsLarge, float>), dim3(batch), dim3(kThreadsLarge), 0, s,
                               sequences, seq_stride, a, angles_stride, c, rn, an, atoms_stride, status);
        else
            hipLaunchKernelGGL((a2c_fwd<kThreadsSmall, float>), dim3(batch), dim3(kThreadsSmall), 0, s,
                               sequences, seq_stride, a, angles_stride, c, rn, an, atoms_stride, status);
    } else {
        const double* a = static_cast<const double*>(angles);
        double* c = static_cast<double*>(coords);
        if (large)
            hipLaunchKernelGGL((a2c_fwd<kThreadsLarge, double>), dim3(batch), dim3(kThreadsLarge), 0, s,
                               sequences, seq_stride, a, angles_stride, c, rn, an, atoms_stride, status);
        else
            hipLaunchKernelGGL((a2c_fwd<kThreadsSmall, double>), dim3(batch), dim3(kThreadsSmall), 0, s,
                               sequences, seq_stride, a, angles_stride, c, rn, an, atoms_stride, status);
    }
    return hipGetLastError() == hipSuccess ? TPL_A2C_OK : TPL_A2C_ERR_LAUNCH;
}

int tpl_a2c_backward(int dtype, const uint8_t* sequences, int seq_stride, const void* coords,
                     const void* grad_coords, int atoms_stride, void* grad_angles, int angles_stride,
                     uint32_t* status, int batch, tpl_a2c_stream_t stream) {
    const int rc = check(dtype, batch, seq_stride, angles_stride, atoms_stride);
    if (rc != TPL_A2C_OK) return rc;
    if (batch == 0 || angles_stride == 0) return TPL_A2C_OK;
    if (!sequences || !grad_angles || (atoms_stride > 0 && (!coords || !grad_coords)))
        return TPL_A2C_ERR_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool large = use_large_bwd(angles_stride);
    if (dtype == TPL_A2C_F32) {
        const float *x = static_cast<const float*>(coords), *g = static_cast<const float*>(grad_coords);
        float* ga = static_cast<float*>(grad_angles);
        if (large)
            hipLaunchKernelGGL((a2c_bwd<kThreadsLarge, float>), dim3(batch), dim3(kThreadsLarge), 0, s,
                               sequences, seq_stride, x, g, atoms_stride, ga, angles_stride, status);
        else
            hipLaunchKernelGGL((a2c_bwd<kThreadsSmall, float>), dim3(batch), dim3(kThreadsSmall), 0, s,
                               sequences, seq_stride, x, g, atoms_stride, ga, angles_stride, status);
    } else {
        const double *x = static_cast<const double*>(coords), *g = static_cast<const double*>(grad_coords);
        double* ga = static_cast<double*>(grad_angles);
        if (large)
            hipLaunchKernelGGL((a2c_bwd<kThreadsLarge, double>), dim3(batch), dim3(kThreadsLarge), 0, s,
                               sequences, seq_stride, x, g, atoms_stride, ga, angles_stride, status);
        else
            hipLaunchKernelGGL((a2c_bwd<kThreadsSmall, double>), dim3(batch), dim3(kThreadsSmall), 0, s,
                               sequences, seq_stride, x, g, atoms_stride, ga, angles_stride, status);
    }
    return hipGetLastError() == hipSuccess ? TPL_A2C_OK : TPL_A2C_ERR_LAUNCH;
}

}  // extern "C"
