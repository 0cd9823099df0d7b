// a2b_cpu.cpp -- Angles2Backbone on the host, FP64 (include/tpl_a2b_cpu.h).
//
// The reference's CPU path (Layers/ReducedModel/cBackboneProteinCPUKernels.cpp)
// builds every atom's 4x4 frame with heap-allocated matrices, then an
// O(L^2) derivative tensor.  Here, per chain:
//   * forward: frame f <- f * Ry(beta) Tx(d) Rx(theta) as the structured 3x4
//     update (27 operations), atom i at the translation of f;
//   * backward: one reverse sweep over the atoms with the suffix sums
//     T = sum_{i>=j} x_i x g_i and F = sum_{i>=j} g_i, and
//     dL/dtheta_j = u_j . (T - x_j x F), u_j = (x_j - x_{j-1}) / d_j.
// Chains are independent and go to a pool of std::threads, one chain at a
// time from a shared counter.
#include "tpl_a2b.h"
#include "tpl_a2b_cpu.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

// Bond lengths and beta = pi - bond angle (cBackboneProteinCPUKernels.cpp:6-12);
// index 0: N->CA (phi), 1: CA->C (psi), 2: C->N (omega).
struct Bond {
    double cb, sb, d;
};
const Bond kBond[3] = {{std::cos(M_PI - 1.9391), std::sin(M_PI - 1.9391), 1.460},
                       {std::cos(M_PI - 2.061), std::sin(M_PI - 2.061), 1.525},
                       {std::cos(M_PI - 2.1186), std::sin(M_PI - 2.1186), 1.330}};

// f (3x4 by columns: R e_x, R e_y, R e_z, t) <- f * B(theta, beta_k, d_k).
inline void frame_step(double f[12], int k, double c, double s) {
    const Bond& b = kBond[k];
    for (int i = 0; i < 3; ++i) {
        const double a0 = f[i], a1 = f[3 + i], a2 = f[6 + i];
        const double n0 = b.cb * a0 - b.sb * a2;
        const double p2 = b.sb * a0 + b.cb * a2;
        f[i] = n0;
        f[3 + i] = c * a1 + s * p2;
        f[6 + i] = c * p2 - s * a1;
        f[9 + i] += b.d * n0;
    }
}

inline int clamp_len(int n, int L, std::atomic<uint32_t>& st) {
    if (n < 0 || n > L) {
        st.fetch_or(TPL_A2B_STATUS_LENGTH);
        return n < 0 ? 0 : L;
    }
    return n;
}

// Atoms N_0, CA_0, C_0, N_1, ...: atom i >= 1 is placed by angle row
// (i-1) % 3, residue (i-1) / 3 (cBackboneProteinCPUKernels.cpp:36-58).
bool chain_forward(const double* ang, double* x, int L, int n) {
    double f[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
    bool bad = false;
    if (n > 0) x[0] = x[1] = x[2] = 0.0;
    for (int i = 1; i < 3 * n; ++i) {
        const int k = (i - 1) % 3, r = (i - 1) / 3;
        const double th = ang[(size_t)k * L + r];
        frame_step(f, k, std::cos(th), std::sin(th));
        for (int c = 0; c < 3; ++c) {
            x[3 * i + c] = f[9 + c];
            bad |= std::isnan(f[9 + c]);
        }
    }
    std::fill(x + 9 * (size_t)n, x + 9 * (size_t)L, 0.0);
    return bad;
}

bool chain_backward(const double* x, const double* g, double* ga, int L, int n) {
    std::fill(ga, ga + 3 * (size_t)L, 0.0);
    double T[3] = {0, 0, 0}, F[3] = {0, 0, 0};
    bool bad = false;
    for (int j = 3 * n - 1; j >= 1; --j) {
        const double* xj = x + 3 * j;
        const double* gj = g + 3 * j;
        T[0] += xj[1] * gj[2] - xj[2] * gj[1];
        T[1] += xj[2] * gj[0] - xj[0] * gj[2];
        T[2] += xj[0] * gj[1] - xj[1] * gj[0];
        F[0] += gj[0];
        F[1] += gj[1];
        F[2] += gj[2];
        const int k = (j - 1) % 3, r = (j - 1) / 3;
        if (r == n - 1 && k > 0) continue;  // psi, omega of the last residue move no atom
        const double inv = 1.0 / kBond[k].d;
        const double u0 = (xj[0] - xj[-3]) * inv, u1 = (xj[1] - xj[-2]) * inv,
                     u2 = (xj[2] - xj[-1]) * inv;
        const double v = u0 * (T[0] - (xj[1] * F[2] - xj[2] * F[1])) +
                         u1 * (T[1] - (xj[2] * F[0] - xj[0] * F[2])) +
                         u2 * (T[2] - (xj[0] * F[1] - xj[1] * F[0]));
        ga[(size_t)k * L + r] = v;
        bad |= std::isnan(v);
    }
    return bad;
}

// Runs fn(b) for b in [0, batch) on `threads` threads (0: hardware threads).
template <typename Fn>
void for_chains(int batch, int threads, Fn fn) {
    int nt = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
    nt = std::max(1, std::min(nt, batch));
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int b = next.fetch_add(1); b < batch; b = next.fetch_add(1)) fn(b);
    };
    if (nt == 1) {
        work();
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(nt - 1);
    for (int t = 1; t < nt; ++t) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
}

int check(int batch, int L) {
    if (batch < 0 || L < 0) return TPL_A2B_ERR_ARG;
    if ((int64_t)9 * L > INT32_MAX || (int64_t)batch * L > INT32_MAX) return TPL_A2B_ERR_SIZE;
    return TPL_A2B_OK;
}

}  // namespace

extern "C" {

int tpl_a2b_cpu_forward(const double* angles, double* coords, const int32_t* lengths,
                        uint32_t* status, int batch, int angles_stride, int threads) {
    const int rc = check(batch, angles_stride);
    if (rc) return rc;
    if (batch == 0 || angles_stride == 0) return TPL_A2B_OK;
    if (!angles || !coords || !lengths) return TPL_A2B_ERR_ARG;
    const int L = angles_stride;
    std::atomic<uint32_t> st{0};
    for_chains(batch, threads, [&](int b) {
        const int n = clamp_len(lengths[b], L, st);
        if (chain_forward(angles + (size_t)b * 3 * L, coords + (size_t)b * 9 * L, L, n))
            st.fetch_or(TPL_A2B_STATUS_NAN);
    });
    if (status) *status |= st.load();
    return TPL_A2B_OK;
}

int tpl_a2b_cpu_backward(double* grad_angles, const double* grad_coords, const double* coords,
                         const int32_t* lengths, uint32_t* status, int batch, int angles_stride,
                         int threads) {
    const int rc = check(batch, angles_stride);
    if (rc) return rc;
    if (batch == 0 || angles_stride == 0) return TPL_A2B_OK;
    if (!grad_angles || !grad_coords || !coords || !lengths) return TPL_A2B_ERR_ARG;
    const int L = angles_stride;
    std::atomic<uint32_t> st{0};
    for_chains(batch, threads, [&](int b) {
        const int n = clamp_len(lengths[b], L, st);
        if (chain_backward(coords + (size_t)b * 9 * L, grad_coords + (size_t)b * 9 * L,
                           grad_angles + (size_t)b * 3 * L, L, n))
            st.fetch_or(TPL_A2B_STATUS_NAN);
    });
    if (status) *status |= st.load();
    return TPL_A2B_OK;
}

}  // extern "C"
