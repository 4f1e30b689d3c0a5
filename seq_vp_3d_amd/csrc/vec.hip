// Free-vector algebra of the inner solver on the device (gfx950): compensated dot products for the
// BB2 proposal, slopes and norms (poisson_gradient.py:955-990, 1263, 1307-1315), and the theta <->
// node-field injection/restriction (boundary.py:8-17). HBM bound: one streaming pass per call.
#include "q1_common.h"
#include "abi_util.h"

namespace seqvp {

constexpr int VB = 256, VGRID = 1024;

struct Pairs {
  const double* x[4];
  const double* y[4];
};

template <int NP>
__global__ void __launch_bounds__(VB) dots_kernel(Pairs p, int64_t n, double* __restrict__ partials) {
  DD s[4] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  const int64_t stride = (int64_t)gridDim.x * VB;
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += stride) {
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      // exact product split (two-prod via fma) keeps the dot product compensated
      const double a = p.x[q][i], b = p.y[q][i];
      const double prod = a * b;
      const double err = fma(a, b, -prod);
      dd_add(s[q], prod);
      s[q].lo += err;
    }
  }
  __shared__ double sh[VB / 64][8];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NP; ++q) s[q] = dd_wave_sum(s[q]);
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      sh[wave][2 * q] = s[q].hi;
      sh[wave][2 * q + 1] = s[q].lo;
    }
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      DD t{sh[0][2 * q], sh[0][2 * q + 1]};
      for (int w = 1; w < VB / 64; ++w) dd_add(t, DD{sh[w][2 * q], sh[w][2 * q + 1]});
      partials[(int64_t)blockIdx.x * 8 + 2 * q] = t.hi;
      partials[(int64_t)blockIdx.x * 8 + 2 * q + 1] = t.lo;
    }
}

__global__ void __launch_bounds__(VB) dots_reduce_kernel(const double* __restrict__ partials, int nblocks, int np,
                                                         double* __restrict__ out) {
  // one wave per pair, fixed order
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  if (q >= np) return;
  DD s{0.0, 0.0};
  for (int b = lane; b < nblocks; b += 64) dd_add(s, DD{partials[b * 8 + 2 * q], partials[b * 8 + 2 * q + 1]});
  s = dd_wave_sum(s);
  if (lane == 0) out[q] = s.hi + s.lo;
}

// theta (3, mz, my, mx) component-major -> node field (nz, ny, nx, 3), zero boundary.
__global__ void __launch_bounds__(VB) inject_kernel(const double* __restrict__ theta, int64_t nx, int64_t ny,
                                                    int64_t nz, double* __restrict__ u) {
  const int64_t nn = nx * ny * nz;
  const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
  if (i >= nn) return;
  const int64_t x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
  const bool inner = x > 0 && x < nx - 1 && y > 0 && y < ny - 1 && z > 0 && z < nz - 1;
  const int64_t mx = nx - 2, my = ny - 2, mz = nz - 2, F = mx * my * mz;
  const int64_t f = ((z - 1) * my + (y - 1)) * mx + (x - 1);
#pragma unroll
  for (int c = 0; c < 3; ++c) u[3 * i + c] = inner ? theta[c * F + f] : 0.0;
}

// node field -> theta (restrict to free nodes, component-major).
__global__ void __launch_bounds__(VB) restrict_kernel(const double* __restrict__ u, int64_t nx, int64_t ny, int64_t nz,
                                                      double* __restrict__ theta) {
  const int64_t mx = nx - 2, my = ny - 2, mz = nz - 2, F = mx * my * mz;
  const int64_t f = (int64_t)blockIdx.x * VB + threadIdx.x;
  if (f >= F) return;
  const int64_t x = f % mx + 1, y = (f / mx) % my + 1, z = f / (mx * my) + 1;
  const int64_t i = (z * ny + y) * nx + x;
#pragma unroll
  for (int c = 0; c < 3; ++c) theta[c * F + f] = u[3 * i + c];
}

// u_trial = inject(theta + tau p): the candidate map of the line search (poisson_gradient.py:1403-1437)
// written straight into the node layout, with the candidate theta itself.
__global__ void __launch_bounds__(VB) axpy_inject_kernel(const double* __restrict__ theta, const double* __restrict__ p,
                                                         double tau, int64_t nx, int64_t ny, int64_t nz,
                                                         double* __restrict__ theta_out, double* __restrict__ u) {
#pragma clang fp contract(off)
  const int64_t nn = nx * ny * nz;
  const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
  if (i >= nn) return;
  const int64_t x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
  const bool inner = x > 0 && x < nx - 1 && y > 0 && y < ny - 1 && z > 0 && z < nz - 1;
  const int64_t mx = nx - 2, my = ny - 2, mz = nz - 2, F = mx * my * mz;
  const int64_t f = ((z - 1) * my + (y - 1)) * mx + (x - 1);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double v = 0.0;
    if (inner) {
      v = theta[c * F + f] + tau * p[c * F + f];   // product and sum rounded separately (no contraction)
      if (theta_out) theta_out[c * F + f] = v;
    }
    u[3 * i + c] = v;
  }
}

// PG metric equilibration scale on the free vector (poisson_gradient.py:472-479, 507-515;
// jd_affine.py:228-242): s = sqrt(m)/sqrt(d) * r with r = min(1, max(J_a, floor)/act) from the node JD
// measure (node == NULL: r = 1). bad[0] counts entries with d <= 0, nonfinite d or nonfinite J_a.
__global__ void __launch_bounds__(VB) metric_scale_kernel(const double* __restrict__ diag,
                                                          const double* __restrict__ node, double sqrt_m, double act,
                                                          double floor_, int64_t nx, int64_t ny, int64_t nz,
                                                          double* __restrict__ scale,
                                                          unsigned long long* __restrict__ bad) {
  const int64_t mx = nx - 2, my = ny - 2, mz = nz - 2, F = mx * my * mz;
  const int64_t f = (int64_t)blockIdx.x * VB + threadIdx.x;
  int nbad = 0;
  if (f < F) {
    double r = 1.0;
    if (node) {
      const int64_t x = f % mx + 1, y = (f / mx) % my + 1, z = f / (mx * my) + 1;
      const double ja = node[(z * ny + y) * nx + x];
      if (!isfinite(ja)) nbad = 1;
      r = fmin(1.0, fmax(ja, floor_) / act);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double d = diag[c * F + f];
      if (!(d > 0.0) || !isfinite(d)) nbad = 1;
      scale[c * F + f] = (sqrt_m / sqrt(d)) * r;
    }
  }
  __shared__ int sh[VB / 64];
  int t = nbad;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) t += __shfl_xor(t, m, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int tot = sh[0] + sh[1] + sh[2] + sh[3];
    if (tot) atomicAdd(bad, (unsigned long long)tot);
  }
}

}  // namespace seqvp

using namespace seqvp;

size_t seqvp_vec_workspace_bytes(int64_t n) {
  (void)n;
  return (size_t)VGRID * 8 * sizeof(double);
}

extern "C" int seqvp_dots(const double* const* xs, const double* const* ys, int npairs, int64_t n, double* out,
                          void* ws, size_t ws_bytes, void* stream) {
  if (!xs || !ys || !out || npairs < 1 || npairs > 4 || n < 0) return SEQVP_EINVAL;
  if (!ws || ws_bytes < seqvp_vec_workspace_bytes(n)) return SEQVP_EWORKSPACE;
  Pairs p{};
  for (int q = 0; q < npairs; ++q) {
    if (!xs[q] || !ys[q]) return SEQVP_EINVAL;
    p.x[q] = xs[q];
    p.y[q] = ys[q];
  }
  hipStream_t s = (hipStream_t)stream;
  double* part = (double*)ws;
  switch (npairs) {
    case 1: hipLaunchKernelGGL((dots_kernel<1>), dim3(VGRID), dim3(VB), 0, s, p, n, part); break;
    case 2: hipLaunchKernelGGL((dots_kernel<2>), dim3(VGRID), dim3(VB), 0, s, p, n, part); break;
    case 3: hipLaunchKernelGGL((dots_kernel<3>), dim3(VGRID), dim3(VB), 0, s, p, n, part); break;
    default: hipLaunchKernelGGL((dots_kernel<4>), dim3(VGRID), dim3(VB), 0, s, p, n, part); break;
  }
  hipLaunchKernelGGL(dots_reduce_kernel, dim3(1), dim3(VB), 0, s, part, VGRID, npairs, out);
  return launch_status();
}

extern "C" int seqvp_inject(const double* theta, int64_t nx, int64_t ny, int64_t nz, double* u, void* stream) {
  if (!theta || !u || nx < 3 || ny < 3 || nz < 3) return SEQVP_EINVAL;
  const int64_t nn = nx * ny * nz;
  hipLaunchKernelGGL(inject_kernel, dim3((nn + VB - 1) / VB), dim3(VB), 0, (hipStream_t)stream, theta, nx, ny, nz, u);
  return launch_status();
}

extern "C" int seqvp_restrict(const double* u, int64_t nx, int64_t ny, int64_t nz, double* theta, void* stream) {
  if (!theta || !u || nx < 3 || ny < 3 || nz < 3) return SEQVP_EINVAL;
  const int64_t F = (nx - 2) * (ny - 2) * (nz - 2);
  hipLaunchKernelGGL(restrict_kernel, dim3((F + VB - 1) / VB), dim3(VB), 0, (hipStream_t)stream, u, nx, ny, nz, theta);
  return launch_status();
}

extern "C" int seqvp_axpy_inject(const double* theta, const double* p, double tau, int64_t nx, int64_t ny, int64_t nz,
                                 double* theta_out, double* u, void* stream) {
  if (!theta || !p || !u || nx < 3 || ny < 3 || nz < 3) return SEQVP_EINVAL;
  const int64_t nn = nx * ny * nz;
  hipLaunchKernelGGL(axpy_inject_kernel, dim3((nn + VB - 1) / VB), dim3(VB), 0, (hipStream_t)stream, theta, p, tau, nx,
                     ny, nz, theta_out, u);
  return launch_status();
}

extern "C" int seqvp_metric_scale(const double* diag, const double* node, double m, double act, double floor_,
                                  int64_t nx, int64_t ny, int64_t nz, double* scale, int64_t* bad, void* stream) {
  if (!diag || !scale || !bad || nx < 3 || ny < 3 || nz < 3 || !(m > 0.0) || (node && !(act > 0.0))) return SEQVP_EINVAL;
  const int64_t F = (nx - 2) * (ny - 2) * (nz - 2);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(copy_or_zero_kernel, dim3(1), dim3(256), 0, s, nullptr, (double*)bad, 1);
  hipLaunchKernelGGL(metric_scale_kernel, dim3((F + VB - 1) / VB), dim3(VB), 0, s, diag, node, sqrt(m), act, floor_, nx,
                     ny, nz, scale, (unsigned long long*)bad);
  return launch_status();
}
