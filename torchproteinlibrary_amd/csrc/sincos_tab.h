// sincos_tab.h -- branch-free FP64 sin and cos from a 64-entry table, for
// device code (the Angles2Backbone and Angles2Coords kernels).  Everything is
// in an anonymous namespace: each translation unit keeps its own copy of the
// constant table.
#pragma once

#include <hip/hip_runtime.h>

namespace {

// sin and cos of x in FP64, table-based.  x = k*(pi/32) + d with pi/32 split
// into three parts (two of 33 significant bits, so k*P1 and k*P2 are exact for
// |k| < 2^20, plus the rounded remainder), |d| <= pi/64.  Then
//   sin x = s_j cos d + c_j sin d,   cos x = c_j cos d - s_j sin d,
// with (c_j, s_j) = (cos, sin)(2 pi j / 64), j = k mod 64, from a table, and
// Taylor series for sin d (to d^7) and cos d (to d^8), truncation < 5e-18.
// Table entries are the exact values rounded to double.  Agreement with libm
// is within 2.2e-16 for |x| < 1e4; the reduction stays exact to |x| < 1e5.
constexpr double kTabInvStep = 0x1.45f306dc9c883p+3;  // 32/pi
constexpr double kTabStep1 = 0x1.921fb54400000p-4;
constexpr double kTabStep2 = 0x1.0b4611a600000p-38;
constexpr double kTabStep3 = 0x1.3198a2e037073p-73;
// cos, sin of 2*pi*j/64 for j = 0..63, interleaved.
__constant__ double kTabCS[128] = {
    0x1.0000000000000p+0, 0x0.0p+0,
    0x1.fd88da3d12526p-1, 0x1.917a6bc29b42cp-4,
    0x1.f6297cff75cb0p-1, 0x1.8f8b83c69a60bp-3,
    0x1.e9f4156c62ddap-1, 0x1.294062ed59f06p-2,
    0x1.d906bcf328d46p-1, 0x1.87de2a6aea963p-2,
    0x1.c38b2f180bdb1p-1, 0x1.e2b5d3806f63bp-2,
    0x1.a9b66290ea1a3p-1, 0x1.1c73b39ae68c8p-1,
    0x1.8bc806b151741p-1, 0x1.44cf325091dd6p-1,
    0x1.6a09e667f3bcdp-1, 0x1.6a09e667f3bcdp-1,
    0x1.44cf325091dd6p-1, 0x1.8bc806b151741p-1,
    0x1.1c73b39ae68c8p-1, 0x1.a9b66290ea1a3p-1,
    0x1.e2b5d3806f63bp-2, 0x1.c38b2f180bdb1p-1,
    0x1.87de2a6aea963p-2, 0x1.d906bcf328d46p-1,
    0x1.294062ed59f06p-2, 0x1.e9f4156c62ddap-1,
    0x1.8f8b83c69a60bp-3, 0x1.f6297cff75cb0p-1,
    0x1.917a6bc29b42cp-4, 0x1.fd88da3d12526p-1,
    -0x1.2ff3e0ec6474dp-192, 0x1.0000000000000p+0,
    -0x1.917a6bc29b42cp-4, 0x1.fd88da3d12526p-1,
    -0x1.8f8b83c69a60bp-3, 0x1.f6297cff75cb0p-1,
    -0x1.294062ed59f06p-2, 0x1.e9f4156c62ddap-1,
    -0x1.87de2a6aea963p-2, 0x1.d906bcf328d46p-1,
    -0x1.e2b5d3806f63bp-2, 0x1.c38b2f180bdb1p-1,
    -0x1.1c73b39ae68c8p-1, 0x1.a9b66290ea1a3p-1,
    -0x1.44cf325091dd6p-1, 0x1.8bc806b151741p-1,
    -0x1.6a09e667f3bcdp-1, 0x1.6a09e667f3bcdp-1,
    -0x1.8bc806b151741p-1, 0x1.44cf325091dd6p-1,
    -0x1.a9b66290ea1a3p-1, 0x1.1c73b39ae68c8p-1,
    -0x1.c38b2f180bdb1p-1, 0x1.e2b5d3806f63bp-2,
    -0x1.d906bcf328d46p-1, 0x1.87de2a6aea963p-2,
    -0x1.e9f4156c62ddap-1, 0x1.294062ed59f06p-2,
    -0x1.f6297cff75cb0p-1, 0x1.8f8b83c69a60bp-3,
    -0x1.fd88da3d12526p-1, 0x1.917a6bc29b42cp-4,
    -0x1.0000000000000p+0, 0x1.08469d45fa26fp-186,
    -0x1.fd88da3d12526p-1, -0x1.917a6bc29b42cp-4,
    -0x1.f6297cff75cb0p-1, -0x1.8f8b83c69a60bp-3,
    -0x1.e9f4156c62ddap-1, -0x1.294062ed59f06p-2,
    -0x1.d906bcf328d46p-1, -0x1.87de2a6aea963p-2,
    -0x1.c38b2f180bdb1p-1, -0x1.e2b5d3806f63bp-2,
    -0x1.a9b66290ea1a3p-1, -0x1.1c73b39ae68c8p-1,
    -0x1.8bc806b151741p-1, -0x1.44cf325091dd6p-1,
    -0x1.6a09e667f3bcdp-1, -0x1.6a09e667f3bcdp-1,
    -0x1.44cf325091dd6p-1, -0x1.8bc806b151741p-1,
    -0x1.1c73b39ae68c8p-1, -0x1.a9b66290ea1a3p-1,
    -0x1.e2b5d3806f63bp-2, -0x1.c38b2f180bdb1p-1,
    -0x1.87de2a6aea963p-2, -0x1.d906bcf328d46p-1,
    -0x1.294062ed59f06p-2, -0x1.e9f4156c62ddap-1,
    -0x1.8f8b83c69a60bp-3, -0x1.f6297cff75cb0p-1,
    -0x1.917a6bc29b42cp-4, -0x1.fd88da3d12526p-1,
    -0x1.2f78b1bbc4ef2p-184, -0x1.0000000000000p+0,
    0x1.917a6bc29b42cp-4, -0x1.fd88da3d12526p-1,
    0x1.8f8b83c69a60bp-3, -0x1.f6297cff75cb0p-1,
    0x1.294062ed59f06p-2, -0x1.e9f4156c62ddap-1,
    0x1.87de2a6aea963p-2, -0x1.d906bcf328d46p-1,
    0x1.e2b5d3806f63bp-2, -0x1.c38b2f180bdb1p-1,
    0x1.1c73b39ae68c8p-1, -0x1.a9b66290ea1a3p-1,
    0x1.44cf325091dd6p-1, -0x1.8bc806b151741p-1,
    0x1.6a09e667f3bcdp-1, -0x1.6a09e667f3bcdp-1,
    0x1.8bc806b151741p-1, -0x1.44cf325091dd6p-1,
    0x1.a9b66290ea1a3p-1, -0x1.1c73b39ae68c8p-1,
    0x1.c38b2f180bdb1p-1, -0x1.e2b5d3806f63bp-2,
    0x1.d906bcf328d46p-1, -0x1.87de2a6aea963p-2,
    0x1.e9f4156c62ddap-1, -0x1.294062ed59f06p-2,
    0x1.f6297cff75cb0p-1, -0x1.8f8b83c69a60bp-3,
    0x1.fd88da3d12526p-1, -0x1.917a6bc29b42cp-4,
};
constexpr double kS3 = -0x1.5555555555555p-3, kS5 = 0x1.1111111111111p-7,
                 kS7 = -0x1.a01a01a01a01ap-13;
constexpr double kC2 = -0x1.0000000000000p-1, kC4 = 0x1.5555555555555p-5,
                 kC6 = -0x1.6c16c16c16c17p-10, kC8 = 0x1.a01a01a01a01ap-16;

// tab: the 128 doubles of kTabCS (in LDS).  Branch-free.  The reduction is
// exact for |k| < 2^20, i.e. |x| < 1e5 rad; beyond that it loses bits
// gradually (2e-10 at |x| = 1e6).  The table index stays in [0, 63] for every
// input, including NaN and inf (which give NaN).
__device__ __forceinline__ void sincos_tab(double x, const double* tab, double* s_out,
                                           double* c_out) {
    const double k = rint(x * kTabInvStep);
    double d = fma(-k, kTabStep1, x);
    d = fma(-k, kTabStep2, d);
    d = fma(-k, kTabStep3, d);
    const int j = static_cast<int>(k) & 63;
    const double cj = tab[2 * j], sj = tab[2 * j + 1];
    const double d2 = d * d;
    const double sd = fma(d * d2, fma(d2, fma(d2, kS7, kS5), kS3), d);
    const double cd = fma(d2, fma(d2, fma(d2, fma(d2, kC8, kC6), kC4), kC2), 1.0);
    *s_out = fma(sj, cd, cj * sd);
    *c_out = fma(cj, cd, -(sj * sd));
}

}  // namespace
