// T75 floating-point Bernstein prescreen of the full-cell JD certificate (gfx950): the kernels over a cell
// list and over every cell. The per-cell arithmetic is prescreen_vertices (prescreen_cell.h), a bitwise
// restatement of jd_curl/three_d/certificate_prescreen.py:60-121.
//
// One cell per lane; the 27 closed-point determinants stay in registers. ~6.5 kflop per cell, no
// FMA: bound by FP64 VALU issue (half the FMA peak).
#include "q1_common.h"
#include "prescreen_cell.h"
#include "abi_util.h"

#pragma clang fp contract(off)

namespace seqvp {

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
  return prescreen_vertices(uv, D);
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
    const bool ok = prescreen_margin(v, delta) > 0.0;
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
