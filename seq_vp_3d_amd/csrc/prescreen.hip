// T75 floating-point Bernstein prescreen of the full-cell JD certificate (gfx950).
//
// Restates jd_curl/three_d/certificate_prescreen.py:60-121 operation for operation: every +, -, *
// rounded separately in binary64, in the same order, with no contraction (this translation unit is
// compiled with -ffp-contract=off and the pragma below), so each cell's value is bitwise the CPU
// kernel's and every proof decision is identical. The bound B is the reference's a-priori bound on
// |beta_fl - beta_exact|; min beta_fl - B > delta proves J > delta on the whole closed cell.
//
// One cell per lane; the 27 closed-point determinants stay in registers. ~6.5 kflop per cell, no
// FMA: bound by FP64 VALU issue (half the FMA peak).
#include "q1_common.h"
#include "abi_util.h"

#pragma clang fp contract(off)

namespace seqvp {

// gamma_n = n u / (1 - n u), u = 2^-53 (certificate_prescreen.py:46-55).
__device__ constexpr double UNIT = 0x1p-53;
__device__ constexpr double ETA = 1e-290;
__device__ constexpr double SQRT27_UP = 5.196152422706633;

__device__ __forceinline__ double gamma_n(double n) { return n * UNIT / (1.0 - n * UNIT); }

__device__ double prescreen_cell(const double* __restrict__ u, const double* __restrict__ D, int64_t c,
                                 const Dims& d) {
  const int64_t x0 = c % d.cx;
  const int64_t y0 = (c / d.cx) % d.cy;
  const int64_t z0 = c / (d.cx * d.cy);
  const int64_t base = (z0 * d.ny + y0) * d.nx + x0;
  double uv[8][3];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int64_t n = base + (a & 1) + ((a >> 1) & 1) * d.nx + ((a >> 2) & 1) * d.nx * d.ny;
    uv[a][0] = u[3 * n];
    uv[a][1] = u[3 * n + 1];
    uv[a][2] = u[3 * n + 2];
  }
  const double g12 = gamma_n(12.0), g6 = gamma_n(6.0), g3 = gamma_n(3.0);
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

__global__ void __launch_bounds__(256) prescreen_kernel(const double* __restrict__ u, const double* __restrict__ D,
                                                        const int64_t* __restrict__ cells, int64_t ncells, Dims d,
                                                        double* __restrict__ out) {
  const int64_t i = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (i >= ncells) return;
  const int64_t c = cells ? cells[i] : i;
  if (c < 0 || c >= d.ncell) {
    out[i] = __builtin_nan("");
    return;
  }
  out[i] = prescreen_cell(u, D, c, d);
}

// flags[c] = proven (the certificate's acceptance rule); per-block counts of unproven cells summed by one atomic per block (integer
// adds commute, so the count is exact and deterministic).
__global__ void __launch_bounds__(256) prescreen_flags_kernel(const double* __restrict__ u,
                                                              const double* __restrict__ D, double delta, Dims d,
                                                              uint8_t* __restrict__ flags,
                                                              unsigned long long* __restrict__ count) {
  const int64_t c = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  int unproven = 0;
  if (c < d.ncell) {
    const double v = prescreen_cell(u, D, c, d);
    // feasibility.py:677-679: proven iff (lower if delta == 0 else nextafter(lower - delta, -inf)) > 0;
    // NaN is never proven.
    const bool ok = delta == 0.0 ? v > 0.0 : nextafter(v - delta, -__builtin_inf()) > 0.0;
    flags[c] = ok ? 1 : 0;
    unproven = ok ? 0 : 1;
  }
  __shared__ int sh[4];
  int t = unproven;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) t += __shfl_xor(t, m, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int tot = sh[0] + sh[1] + sh[2] + sh[3];
    if (tot) atomicAdd(count, (unsigned long long)tot);
  }
}

}  // namespace seqvp

using namespace seqvp;

extern "C" int seqvp_prescreen(const double* u, const double* D, const int64_t* cells, int64_t ncells, int64_t nx,
                               int64_t ny, int64_t nz, double* out, void* stream) {
  if (!u || !D || !out || nx < 2 || ny < 2 || nz < 2 || ncells < 0) return SEQVP_EINVAL;
  const Dims d = make_dims(nx, ny, nz);
  if (!cells && ncells != d.ncell) return SEQVP_EINVAL;
  if (ncells == 0) return SEQVP_OK;
  const int64_t nb = (ncells + 255) / 256;
  hipLaunchKernelGGL(prescreen_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, u, D, cells, ncells, d, out);
  return launch_status();
}

extern "C" int seqvp_prescreen_flags(const double* u, const double* D, double delta, int64_t nx, int64_t ny,
                                     int64_t nz, uint8_t* flags, int64_t* out_count, void* stream) {
  if (!u || !D || !flags || !out_count || nx < 2 || ny < 2 || nz < 2) return SEQVP_EINVAL;
  const Dims d = make_dims(nx, ny, nz);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(copy_or_zero_kernel, dim3(1), dim3(256), 0, s, nullptr, (double*)out_count, 1);
  const int64_t nb = (d.ncell + 255) / 256;
  hipLaunchKernelGGL(prescreen_flags_kernel, dim3(nb), dim3(256), 0, s, u, D, delta, d, flags,
                     (unsigned long long*)out_count);
  return launch_status();
}
