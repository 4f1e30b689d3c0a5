// T75 floating-point Bernstein prescreen of one cell, shared by the GPU kernels and the CPU certificate.
//
// Restates jd_curl/three_d/certificate_prescreen.py:60-121 operation for operation: every +, -, *
// rounded separately in binary64, in the same order, with no contraction (units including this header
// are compiled with -ffp-contract=off and the pragma below), so each cell's value is bitwise the CPU
// kernel's and every proof decision is identical. The bound B is the reference's a-priori bound on
// |beta_fl - beta_exact|; min beta_fl - B > delta proves J > delta on the whole closed cell.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace seqvp {

// gamma_n = n u / (1 - n u), u = 2^-53 (certificate_prescreen.py:46-55).
__host__ __device__ __forceinline__ double prescreen_gamma(double n) {
  return n * 0x1p-53 / (1.0 - n * 0x1p-53);
}

// Lower bound of min_{cell} J from the 8 vertex displacements uv[a][r] (vertex a = x + 2y + 4z) and the
// closed27 derivative table D[(j*27 + p)*8 + a]; NaN when any value is nonfinite.
__host__ __device__ inline double prescreen_vertices(const double uv[8][3], const double* __restrict__ D) {
  const double ETA = 1e-290;
  const double SQRT27_UP = 5.196152422706633;
  const double g12 = prescreen_gamma(12.0), g6 = prescreen_gamma(6.0), g3 = prescreen_gamma(3.0);
  double b[27];
  double sigma = 0.0, phi = 0.0, big = 0.0;
  for (int p = 0; p < 27; ++p) {
    double F[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double acc = 0.0, mag = 0.0;
#pragma unroll
        for (int a = 0; a < 8; ++a) {
          const double term = uv[a][r] * D[(j * 27 + p) * 8 + a];
          acc += term;
          mag += fabs(term);
        }
        F[r][j] = acc;
        if (mag > sigma) sigma = mag;
      }
      F[r][r] += 1.0;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        if (fabs(F[r][j]) > phi) phi = fabs(F[r][j]);
    const double v = (F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) +
                      F[0][1] * (F[1][2] * F[2][0] - F[1][0] * F[2][2])) +
                     F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]);
    b[p] = v;
    if (fabs(v) > big) big = fabs(v);
  }
  // separable T along x, y, z (index 9iz+3iy+ix): middle entry -> 2 b - (b_l + b_r)/2
#pragma unroll
  for (int pass = 0; pass < 3; ++pass) {
    const int stride = pass == 0 ? 1 : (pass == 1 ? 3 : 9);
#pragma unroll
    for (int p = 0; p < 27; ++p)
      if ((p / stride) % 3 == 1) b[p] = 2.0 * b[p] - 0.5 * (b[p - stride] + b[p + stride]);
  }
  const double e = g12 * (sigma + phi) + ETA;
  const double P = phi + e;
  const double ej = SQRT27_UP * e * (3.0 * P * P + 3.0 * P * e + e * e) + 6.0 * g6 * P * P * P;
  const double bound = 1.01 * (27.0 * ej + 135.0 * g3 * big) + ETA;
  double low = __builtin_inf();
  bool finite = true;
  for (int p = 0; p < 27; ++p) {
    const double w = b[p];
    if (!(w == w) || w == __builtin_inf() || w == -__builtin_inf())
      finite = false;
    else if (w < low)
      low = w;
  }
  if (!(bound == bound) || bound == __builtin_inf() || !(sigma == sigma) || sigma == __builtin_inf()) finite = false;
  return finite ? low - bound : __builtin_nan("");
}

// feasibility.py:677-679: the prescreen margin (lower if delta == 0 else nextafter(lower - delta, -inf)).
__host__ __device__ __forceinline__ double prescreen_margin(double lower, double delta) {
  if (delta == 0.0) return lower;
  const double x = lower - delta;
  if (!(x == x) || x == -__builtin_inf()) return x;
  if (x == 0.0) return -__builtin_bit_cast(double, 1ull);
  const unsigned long long bits = __builtin_bit_cast(unsigned long long, x);
  return __builtin_bit_cast(double, x > 0.0 ? bits - 1 : bits + 1);
}

}  // namespace seqvp
