// Pyramid resampling on the device, float64 (gfx950): the image pyramid's Gaussian + linear downsampling
// (jd_curl/three_d/nifti.py:128-149: gaussian_filter mode "nearest", truncate 4, then order-1
// map_coordinates at linspace positions, mode "nearest"), the Q1 prolongation of a displacement to a finer
// grid (multiresolution.py:55-84), the boundary taper (preopt.py:617-624) and the conversions between the
// (nz, ny, nx, 3) mm field and the pre-optimiser's (3, D-2, H-2, W-2) voxel-unit interior array.
//
// Each pass is separable along one axis of a C-order (n0, n1, n2, ncomp) array. The linear pass computes
// (1 - w) a + w b with separately rounded operations (this unit is compiled without contraction), which is
// the prolongation's own arithmetic; the downsampling's trilinear map_coordinates is the product of the
// three linear passes (equal up to rounding).
#include <algorithm>

#include "abi_util.h"

#pragma clang fp contract(off)

namespace seqvp {
namespace rs {

static constexpr int RB = 256;

static inline unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>((n + RB - 1) / RB, 4096); }

// out = correlation of in with w[0..2r] along `axis`, replicate padding.
__global__ void __launch_bounds__(RB) gauss_axis_kernel(const double* __restrict__ in, double* __restrict__ out,
                                                        int64_t n0, int64_t n1, int64_t n2, int64_t nc, int axis,
                                                        const double* __restrict__ w, int r) {
  const int64_t total = n0 * n1 * n2 * nc;
  const int64_t len = axis == 0 ? n0 : (axis == 1 ? n1 : n2);
  const int64_t stride = (axis == 0 ? n1 * n2 : (axis == 1 ? n2 : 1)) * nc;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < total; i += (int64_t)gridDim.x * RB) {
    const int64_t pos = (i / stride) % len;
    const int64_t base = i - pos * stride;
    double acc = w[r] * in[i];
    for (int k = 1; k <= r; ++k) {
      const int64_t lo = pos - k < 0 ? 0 : pos - k;
      const int64_t hi = pos + k >= len ? len - 1 : pos + k;
      acc += (in[base + lo * stride] + in[base + hi * stride]) * w[r + k];
    }
    out[i] = acc;
  }
}

// out[.., j, ..] = (1 - w_j) in[.., left_j, ..] + w_j in[.., right_j, ..] along `axis` (output length m).
__global__ void __launch_bounds__(RB) lerp_axis_kernel(const double* __restrict__ in, double* __restrict__ out,
                                                       int64_t n0, int64_t n1, int64_t n2, int64_t nc, int axis,
                                                       int64_t m, const int64_t* __restrict__ left,
                                                       const int64_t* __restrict__ right,
                                                       const double* __restrict__ weight) {
  const int64_t o0 = axis == 0 ? m : n0, o1 = axis == 1 ? m : n1, o2 = axis == 2 ? m : n2;
  const int64_t total = o0 * o1 * o2 * nc;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < total; i += (int64_t)gridDim.x * RB) {
    int64_t t = i;
    const int64_t c = t % nc;
    t /= nc;
    const int64_t i2 = t % o2;
    t /= o2;
    const int64_t i1 = t % o1;
    const int64_t i0 = t / o1;
    const int64_t j = axis == 0 ? i0 : (axis == 1 ? i1 : i2);
    int64_t a0 = i0, a1 = i1, a2 = i2, b0 = i0, b1 = i1, b2 = i2;
    if (axis == 0) { a0 = left[j]; b0 = right[j]; }
    if (axis == 1) { a1 = left[j]; b1 = right[j]; }
    if (axis == 2) { a2 = left[j]; b2 = right[j]; }
    const double wj = weight[j];
    const double va = in[((a0 * n1 + a1) * n2 + a2) * nc + c];
    const double vb = in[((b0 * n1 + b1) * n2 + b2) * nc + c];
    out[i] = (1.0 - wj) * va + wj * vb;
  }
}

// out = ((rz[z] ry[y]) rx[x]) u for a (nz, ny, nx, 3) field.
__global__ void __launch_bounds__(RB) taper_kernel(const double* __restrict__ u, int64_t nz, int64_t ny, int64_t nx,
                                                   const double* __restrict__ rz, const double* __restrict__ ry,
                                                   const double* __restrict__ rx, double* __restrict__ out) {
  const int64_t total = nz * ny * nx * 3;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < total; i += (int64_t)gridDim.x * RB) {
    const int64_t node = i / 3;
    const int64_t x = node % nx, y = (node / nx) % ny, z = node / (nx * ny);
    out[i] = ((rz[z] * ry[y]) * rx[x]) * u[i];
  }
}

__global__ void __launch_bounds__(RB) scale_kernel(const double* __restrict__ u, double a, double* __restrict__ out,
                                                   int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) out[i] = a * u[i];
}

// (nz, ny, nx, 3) mm field -> (3, nz-2, ny-2, nx-2) voxel-unit interior, as T.
template <typename T>
__global__ void __launch_bounds__(RB) to_interior_kernel(const double* __restrict__ u, int64_t nz, int64_t ny,
                                                         int64_t nx, double hx, double hy, double hz,
                                                         T* __restrict__ out) {
  const int64_t Di = nz - 2, Hi = ny - 2, Wi = nx - 2, nint = Di * Hi * Wi;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < 3 * nint; i += (int64_t)gridDim.x * RB) {
    const int64_t c = i / nint, k = i % nint;
    const int64_t x = k % Wi + 1, y = (k / Wi) % Hi + 1, z = k / (Wi * Hi) + 1;
    const double h = c == 0 ? hx : (c == 1 ? hy : hz);
    out[i] = (T)(u[((z * ny + y) * nx + x) * 3 + c] / h);
  }
}

// (3, nz-2, ny-2, nx-2) voxel-unit interior (T) -> (nz, ny, nx, 3) mm field with zero faces.
template <typename T>
__global__ void __launch_bounds__(RB) from_interior_kernel(const T* __restrict__ v, int64_t nz, int64_t ny, int64_t nx,
                                                           double hx, double hy, double hz, double* __restrict__ out) {
  const int64_t total = nz * ny * nx * 3;
  const int64_t Di = nz - 2, Hi = ny - 2, Wi = nx - 2, nint = Di * Hi * Wi;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < total; i += (int64_t)gridDim.x * RB) {
    const int64_t c = i % 3, node = i / 3;
    const int64_t x = node % nx, y = (node / nx) % ny, z = node / (nx * ny);
    double val = 0.0;
    if (x > 0 && x < nx - 1 && y > 0 && y < ny - 1 && z > 0 && z < nz - 1) {
      const double h = c == 0 ? hx : (c == 1 ? hy : hz);
      val = (double)v[c * nint + ((z - 1) * Hi + (y - 1)) * Wi + (x - 1)] * h;
    }
    out[i] = val;
  }
}

}  // namespace rs
}  // namespace seqvp

using namespace seqvp::rs;

extern "C" int seqvp_gauss_axis(const double* in, double* out, int64_t n0, int64_t n1, int64_t n2, int64_t nc,
                                int axis, const double* w, int r, void* stream) {
  if (!in || !out || !w || in == out || n0 < 1 || n1 < 1 || n2 < 1 || nc < 1 || axis < 0 || axis > 2 || r < 0)
    return SEQVP_EINVAL;
  const int64_t total = n0 * n1 * n2 * nc;
  hipLaunchKernelGGL(gauss_axis_kernel, dim3(grid_for(total)), dim3(RB), 0, (hipStream_t)stream, in, out, n0, n1, n2,
                     nc, axis, w, r);
  return launch_status();
}

extern "C" int seqvp_lerp_axis(const double* in, double* out, int64_t n0, int64_t n1, int64_t n2, int64_t nc,
                               int axis, int64_t m, const int64_t* left, const int64_t* right, const double* weight,
                               void* stream) {
  if (!in || !out || !left || !right || !weight || in == out || n0 < 1 || n1 < 1 || n2 < 1 || nc < 1 || axis < 0 ||
      axis > 2 || m < 1)
    return SEQVP_EINVAL;
  const int64_t total = (axis == 0 ? m : n0) * (axis == 1 ? m : n1) * (axis == 2 ? m : n2) * nc;
  hipLaunchKernelGGL(lerp_axis_kernel, dim3(grid_for(total)), dim3(RB), 0, (hipStream_t)stream, in, out, n0, n1, n2,
                     nc, axis, m, left, right, weight);
  return launch_status();
}

extern "C" int seqvp_taper(const double* u, int64_t nx, int64_t ny, int64_t nz, const double* rx, const double* ry,
                           const double* rz, double* out, void* stream) {
  if (!u || !out || !rx || !ry || !rz || nx < 1 || ny < 1 || nz < 1) return SEQVP_EINVAL;
  hipLaunchKernelGGL(taper_kernel, dim3(grid_for(nx * ny * nz * 3)), dim3(RB), 0, (hipStream_t)stream, u, nz, ny, nx,
                     rz, ry, rx, out);
  return launch_status();
}

extern "C" int seqvp_scale(const double* u, double a, double* out, int64_t n, void* stream) {
  if (!u || !out || n < 0) return SEQVP_EINVAL;
  if (n == 0) return SEQVP_OK;
  hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(RB), 0, (hipStream_t)stream, u, a, out, n);
  return launch_status();
}

extern "C" int seqvp_field_to_interior(const double* u, int64_t nx, int64_t ny, int64_t nz, double hx, double hy,
                                       double hz, int fp32, void* out, void* stream) {
  if (!u || !out || nx < 3 || ny < 3 || nz < 3 || !(hx > 0 && hy > 0 && hz > 0)) return SEQVP_EINVAL;
  const unsigned g = grid_for(3 * (nx - 2) * (ny - 2) * (nz - 2));
  if (fp32)
    hipLaunchKernelGGL(to_interior_kernel<float>, dim3(g), dim3(RB), 0, (hipStream_t)stream, u, nz, ny, nx, hx, hy, hz,
                       (float*)out);
  else
    hipLaunchKernelGGL(to_interior_kernel<double>, dim3(g), dim3(RB), 0, (hipStream_t)stream, u, nz, ny, nx, hx, hy,
                       hz, (double*)out);
  return launch_status();
}

extern "C" int seqvp_interior_to_field(const void* v, int64_t nx, int64_t ny, int64_t nz, double hx, double hy,
                                       double hz, int fp32, double* out, void* stream) {
  if (!v || !out || nx < 3 || ny < 3 || nz < 3 || !(hx > 0 && hy > 0 && hz > 0)) return SEQVP_EINVAL;
  const unsigned g = grid_for(3 * nx * ny * nz);
  if (fp32)
    hipLaunchKernelGGL(from_interior_kernel<float>, dim3(g), dim3(RB), 0, (hipStream_t)stream, (const float*)v, nz, ny,
                       nx, hx, hy, hz, out);
  else
    hipLaunchKernelGGL(from_interior_kernel<double>, dim3(g), dim3(RB), 0, (hipStream_t)stream, (const double*)v, nz,
                       ny, nx, hx, hy, hz, out);
  return launch_status();
}
