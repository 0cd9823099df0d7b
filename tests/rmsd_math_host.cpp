// Host build of the device eigen-solve (torchproteinlibrary_amd/csrc/rmsd_math.h)
// for tests/test_rmsd_math.py: the same source the HIP kernel compiles,
// exported with C linkage for ctypes.
#include "../torchproteinlibrary_amd/csrc/rmsd_math.h"

extern "C" {

// R: 3x3 centred correlation (row-major); bound: R2 / 2.
// out: lambda, q0..q3, degenerate flag.
void rmsd_math_eig(const double* R, double bound, double* out) {
    double F[16];
    rmsd_math::f_matrix(R, F);
    const rmsd_math::Eigenpair e = rmsd_math::max_eigenpair(F, bound);
    out[0] = e.lam;
    out[1] = e.q0;
    out[2] = e.q1;
    out[3] = e.q2;
    out[4] = e.q3;
    out[5] = e.degenerate;
}

void rmsd_math_f(const double* R, double* F) { rmsd_math::f_matrix(R, F); }

void rmsd_math_u(const double* q, double* U) {
    const rmsd_math::Eigenpair e{0.0, q[0], q[1], q[2], q[3], 0};
    rmsd_math::u_from_q(e, U);
}

}  // extern "C"
