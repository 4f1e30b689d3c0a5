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
// written straight into the node layout, with the candidate theta itself. tau_dev (when not NULL) holds the step in
// device memory, for captured segments whose step changes between launches.
__global__ void __launch_bounds__(VB) axpy_inject_kernel(const double* __restrict__ theta, const double* __restrict__ p,
                                                         double tau_arg, const double* __restrict__ tau_dev,
                                                         int64_t nx, int64_t ny, int64_t nz,
                                                         double* __restrict__ theta_out, double* __restrict__ u) {
#pragma clang fp contract(off)
  const double tau = tau_dev ? *tau_dev : tau_arg;
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

// BB2 secant products (poisson_gradient.py:955-990): s = theta - theta_old, y = g - g_old,
// iy = p_old - p (each rounded once, as the host subtractions), then compensated s.y, s.s, y.y, y.iy.
__global__ void __launch_bounds__(VB) bb2_kernel(const double* __restrict__ th, const double* __restrict__ th0,
                                                 const double* __restrict__ g, const double* __restrict__ g0,
                                                 const double* __restrict__ p0, const double* __restrict__ p, int64_t n,
                                                 double* __restrict__ partials) {
  DD acc[4] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  const int64_t stride = (int64_t)gridDim.x * VB;
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += stride) {
    const double sv = th[i] - th0[i], yv = g[i] - g0[i], iy = p0[i] - p[i];
    const double pr[4][2] = {{sv, yv}, {sv, sv}, {yv, yv}, {yv, iy}};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double prod = pr[q][0] * pr[q][1];
      const double err = fma(pr[q][0], pr[q][1], -prod);
      dd_add(acc[q], prod);
      acc[q].lo += err;
    }
  }
  __shared__ double sh[VB / 64][8];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 4; ++q) acc[q] = dd_wave_sum(acc[q]);
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      sh[wave][2 * q] = acc[q].hi;
      sh[wave][2 * q + 1] = acc[q].lo;
    }
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      DD t{sh[0][2 * q], sh[0][2 * q + 1]};
      for (int w = 1; w < VB / 64; ++w) dd_add(t, DD{sh[w][2 * q], sh[w][2 * q + 1]});
      partials[(int64_t)blockIdx.x * 8 + 2 * q] = t.hi;
      partials[(int64_t)blockIdx.x * 8 + 2 * q + 1] = t.lo;
    }
}

// Exact linear node-domain step bound (poisson_gradient.py:993-1011): min over free nodes and components of
// (hi - x)/p for p > 0 and (lo - x)/p for p < 0 with x = X + u; +inf if none. Also max |p| (for
// diagnostics). min/max are order independent, so the result equals the host scan's.
__global__ void __launch_bounds__(VB) step_bound_kernel(const double* __restrict__ p, const double* __restrict__ u,
                                                        double ox, double oy, double oz, double hx, double hy,
                                                        double hz, int64_t nx, int64_t ny, int64_t nz,
                                                        double* __restrict__ partials) {
  const int64_t mx = nx - 2, my = ny - 2, mz = nz - 2, F = mx * my * mz;
  const double lo[3] = {ox, oy, oz};
  const double hi[3] = {ox + hx * (double)(nx - 1), oy + hy * (double)(ny - 1), oz + hz * (double)(nz - 1)};
  const double org[3] = {ox, oy, oz}, h[3] = {hx, hy, hz};
  double best = __builtin_inf();
  const int64_t stride = (int64_t)gridDim.x * VB;
  for (int64_t f = (int64_t)blockIdx.x * VB + threadIdx.x; f < F; f += stride) {
    const int64_t x = f % mx + 1, y = (f / mx) % my + 1, z = f / (mx * my) + 1;
    const int64_t node = (z * ny + y) * nx + x;
    const int64_t idx[3] = {x, y, z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double pc = p[c * F + f];
      const double xc = (org[c] + h[c] * (double)idx[c]) + u[3 * node + c];
      if (pc > 0.0) best = fmin(best, (hi[c] - xc) / pc);
      else if (pc < 0.0) best = fmin(best, (lo[c] - xc) / pc);
    }
  }
  __shared__ double sh[VB / 64];
  best = wave_min(best);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = fmin(fmin(sh[0], sh[1]), fmin(sh[2], sh[3]));
}

// max |a - b| and max |b| (step_ratio, common.py: relative step of a candidate).
__global__ void __launch_bounds__(VB) maxabs_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                                    int64_t n, double* __restrict__ partials) {
  double m0 = 0.0, m1 = 0.0;
  const int64_t stride = (int64_t)gridDim.x * VB;
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += stride) {
    const double d = a[i] - b[i];
    m0 = (fabs(d) > m0 || d != d) ? fabs(d) : m0;   // NaN propagates
    m1 = (fabs(b[i]) > m1 || b[i] != b[i]) ? fabs(b[i]) : m1;
  }
  __shared__ double sh[VB / 64][2];
  m0 = wave_max(m0);
  m1 = wave_max(m1);
  if ((threadIdx.x & 63) == 0) {
    sh[threadIdx.x >> 6][0] = m0;
    sh[threadIdx.x >> 6][1] = m1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = fmax(fmax(sh[0][0], sh[1][0]), fmax(sh[2][0], sh[3][0]));
    partials[2 * blockIdx.x + 1] = fmax(fmax(sh[0][1], sh[1][1]), fmax(sh[2][1], sh[3][1]));
  }
}

__global__ void __launch_bounds__(64) fold_min_max_kernel(const double* __restrict__ partials, int nblocks, int mode,
                                                          double* __restrict__ out) {
  // mode 0: min of partials[b]; mode 1: max of partials[2b] and of partials[2b+1]
  const int lane = threadIdx.x;
  if (mode == 0) {
    double m = __builtin_inf();
    for (int b = lane; b < nblocks; b += 64) m = fmin(m, partials[b]);
    m = wave_min(m);
    if (lane == 0) out[0] = m;
  } else {
    double m0 = 0.0, m1 = 0.0;
    for (int b = lane; b < nblocks; b += 64) {
      m0 = fmax(m0, partials[2 * b]);
      m1 = fmax(m1, partials[2 * b + 1]);
    }
    m0 = wave_max(m0);
    m1 = wave_max(m1);
    if (lane == 0) {
      out[0] = m0;
      out[1] = m1;
    }
  }
}

// The scalars of a PG direction in one pass over the free vector (n = 3F entries, component-major): g.p and g.g
// as dots_kernel<2> forms them; with the previous accepted state (th0, g0, p0 not NULL) the BB2 products as
// bb2_kernel forms them and the last step's max |th - th0|, max |th0| as maxabs_kernel does; and the domain step
// bound of p at X + u as step_bound_kernel does. Same per-element operations in the same grid-stride order per
// accumulator, so every value is the bits of its separate kernel. Partials: 16 doubles per block (six
// double-double sums, the bound, the two maxima).
__global__ void __launch_bounds__(VB) direction_scalars_kernel(const double* __restrict__ th,
                                                               const double* __restrict__ th0,
                                                               const double* __restrict__ g,
                                                               const double* __restrict__ g0,
                                                               const double* __restrict__ p0,
                                                               const double* __restrict__ p,
                                                               const double* __restrict__ u, double ox, double oy,
                                                               double oz, double hx, double hy, double hz, int64_t nx,
                                                               int64_t ny, int64_t nz, double* __restrict__ partials) {
  const int64_t mx = nx - 2, my = ny - 2, mz = nz - 2, F = mx * my * mz, n = 3 * F;
  const double lo[3] = {ox, oy, oz};
  const double hi[3] = {ox + hx * (double)(nx - 1), oy + hy * (double)(ny - 1), oz + hz * (double)(nz - 1)};
  const double h[3] = {hx, hy, hz};
  const bool prev = th0 != nullptr;
  DD acc[6] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  double best = __builtin_inf(), m0 = 0.0, m1 = 0.0;
  const int64_t stride = (int64_t)gridDim.x * VB;
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += stride) {
    const double gi = g[i], pi = p[i];
    const double dp[2][2] = {{gi, pi}, {gi, gi}};
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const double prod = dp[q][0] * dp[q][1];
      const double err = fma(dp[q][0], dp[q][1], -prod);
      dd_add(acc[q], prod);
      acc[q].lo += err;
    }
    if (prev) {
      const double t0 = th0[i];
      const double sv = th[i] - t0, yv = gi - g0[i], iy = p0[i] - pi;
      const double pr[4][2] = {{sv, yv}, {sv, sv}, {yv, yv}, {yv, iy}};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double prod = pr[q][0] * pr[q][1];
        const double err = fma(pr[q][0], pr[q][1], -prod);
        dd_add(acc[2 + q], prod);
        acc[2 + q].lo += err;
      }
      m0 = (fabs(sv) > m0 || sv != sv) ? fabs(sv) : m0;   // NaN propagates
      m1 = (fabs(t0) > m1 || t0 != t0) ? fabs(t0) : m1;
    }
    const int64_t c = i / F, f = i - c * F;
    const int64_t x = f % mx + 1, y = (f / mx) % my + 1, z = f / (mx * my) + 1;
    const int64_t idx = c == 0 ? x : (c == 1 ? y : z);
    const double xc = (lo[c] + h[c] * (double)idx) + u[3 * ((z * ny + y) * nx + x) + c];
    if (pi > 0.0) best = fmin(best, (hi[c] - xc) / pi);
    else if (pi < 0.0) best = fmin(best, (lo[c] - xc) / pi);
  }
  __shared__ double sh[VB / 64][16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 6; ++q) acc[q] = dd_wave_sum(acc[q]);
  best = wave_min(best);
  m0 = wave_max(m0);
  m1 = wave_max(m1);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      sh[wave][2 * q] = acc[q].hi;
      sh[wave][2 * q + 1] = acc[q].lo;
    }
    sh[wave][12] = best;
    sh[wave][13] = m0;
    sh[wave][14] = m1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double* o = partials + (int64_t)blockIdx.x * 16;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      DD t{sh[0][2 * q], sh[0][2 * q + 1]};
      for (int w = 1; w < VB / 64; ++w) dd_add(t, DD{sh[w][2 * q], sh[w][2 * q + 1]});
      o[2 * q] = t.hi;
      o[2 * q + 1] = t.lo;
    }
    o[12] = fmin(fmin(sh[0][12], sh[1][12]), fmin(sh[2][12], sh[3][12]));
    o[13] = fmax(fmax(sh[0][13], sh[1][13]), fmax(sh[2][13], sh[3][13]));
    o[14] = fmax(fmax(sh[0][14], sh[1][14]), fmax(sh[2][14], sh[3][14]));
    o[15] = 0.0;
  }
}

// out[0] g.p, out[1] g.g, out[4..7] the BB2 products, out[8] the step bound, out[9..10] the step maxima (each sum
// folded as dots_reduce_kernel folds it, one wave per sum; the bound and maxima as fold_min_max_kernel folds them).
__global__ void __launch_bounds__(512) direction_scalars_reduce_kernel(const double* __restrict__ partials, int nblocks,
                                                                       int prev, double* __restrict__ out) {
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  if (q < 6) {
    if (q >= 2 && !prev) return;
    DD s{0.0, 0.0};
    for (int b = lane; b < nblocks; b += 64) dd_add(s, DD{partials[b * 16 + 2 * q], partials[b * 16 + 2 * q + 1]});
    s = dd_wave_sum(s);
    if (lane == 0) out[q < 2 ? q : q + 2] = s.hi + s.lo;
  } else if (q == 6) {
    double m = __builtin_inf();
    for (int b = lane; b < nblocks; b += 64) m = fmin(m, partials[b * 16 + 12]);
    m = wave_min(m);
    if (lane == 0) out[8] = m;
  } else if (prev) {
    double a = 0.0, c = 0.0;
    for (int b = lane; b < nblocks; b += 64) {
      a = fmax(a, partials[b * 16 + 13]);
      c = fmax(c, partials[b * 16 + 14]);
    }
    a = wave_max(a);
    c = wave_max(c);
    if (lane == 0) {
      out[9] = a;
      out[10] = c;
    }
  }
}

}  // namespace seqvp

using namespace seqvp;

size_t seqvp_vec_workspace_bytes(int64_t n) {
  (void)n;
  return (size_t)VGRID * 16 * sizeof(double);
}

extern "C" int seqvp_direction_scalars(const double* theta, const double* theta_old, const double* g,
                                       const double* g_old, const double* p_old, const double* p, const double* u,
                                       double ox, double oy, double oz, double hx, double hy, double hz, int64_t nx,
                                       int64_t ny, int64_t nz, double* out, void* ws, size_t ws_bytes, void* stream) {
  if (!theta || !g || !p || !u || !out || nx < 3 || ny < 3 || nz < 3) return SEQVP_EINVAL;
  const bool prev = theta_old != nullptr;
  if (prev && (!g_old || !p_old)) return SEQVP_EINVAL;
  if (!ws || ws_bytes < seqvp_vec_workspace_bytes(0)) return SEQVP_EWORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  double* part = (double*)ws;
  hipLaunchKernelGGL(direction_scalars_kernel, dim3(VGRID), dim3(VB), 0, s, theta, theta_old, g, g_old, p_old, p, u,
                     ox, oy, oz, hx, hy, hz, nx, ny, nz, part);
  hipLaunchKernelGGL(direction_scalars_reduce_kernel, dim3(1), dim3(512), 0, s, part, VGRID, prev ? 1 : 0, out);
  return launch_status();
}

namespace seqvp {

// The first attempt's step from the direction scalars (`s` as seqvp_direction_scalars leaves them): the BB2
// proposal with the reference's fallbacks (poisson_gradient.py:955-990), clipped to [step_min, step_max], then
// min(proposal, 0.9 bound) (poisson_gradient.py:1340-1356). Written as the solver's host code computes it,
// operation for operation (IEEE sqrt and division, no contraction), so both give the same bits; min/max follow the
// host's order of arguments for NaN (fmin/fmax return the number).
__global__ void pg_step_kernel(const double* __restrict__ s, int has_prev, double prev_tau, double step_min,
                               double step_max, double initial_step, double* __restrict__ tau) {
#pragma clang fp contract(off)
  auto protect = [&](double v) { return fmin(step_max, fmax(step_min, v)); };
  double proposal;
  if (!has_prev) {
    proposal = protect(initial_step);
  } else {
    const double fallback = protect(2.0 * fmin(prev_tau, step_max / 2.0));
    const double sy = s[4], ns = sqrt(s[5]), ny = sqrt(s[6]), yiy = s[7];
    if (!(isfinite(sy) && isfinite(ns) && isfinite(ny) && isfinite(yiy))) {
      proposal = fallback;
    } else if (sy <= 0.0) {
      proposal = fallback;
    } else {
      const double cosine = (ns > 0.0 && ny > 0.0) ? sy / (ns * ny) : 0.0;
      if (!isfinite(cosine) || cosine <= 1e-12) {
        proposal = fallback;
      } else if (yiy <= 0.0) {
        proposal = fallback;
      } else {
        const double q = sy / yiy;
        proposal = (!isfinite(q) || q <= 0.0) ? fallback : protect(q);
      }
    }
  }
  const double capped = 0.9 * s[8];
  tau[0] = capped < proposal ? capped : proposal;
}

// The trial step capped by the T66 root (poisson_gradient.py:1356-1366): root = out[0] * limit in step units;
// unchanged when the sweep met a nonfinite coefficient (out[2] != 0), found no root, or the root is 0.
__global__ void capped_step_kernel(const double* __restrict__ limit, const double* __restrict__ root_out,
                                   double safety, double* __restrict__ tau) {
#pragma clang fp contract(off)
  const double L = limit[0], rs = root_out[0];
  double t = L;
  if (root_out[2] == 0.0 && isfinite(rs)) {
    const double root = rs * L;
    if (root != 0.0) {
      const double c = safety * root;
      t = c < t ? c : t;
    }
  }
  tau[0] = t;
}

}  // namespace seqvp

extern "C" int seqvp_pg_step(const double* scalars, int has_prev, double prev_tau, double step_min, double step_max,
                             double initial_step, double* tau, void* stream) {
  if (!scalars || !tau) return SEQVP_EINVAL;
  hipLaunchKernelGGL(pg_step_kernel, dim3(1), dim3(1), 0, (hipStream_t)stream, scalars, has_prev, prev_tau,
                     step_min, step_max, initial_step, tau);
  return launch_status();
}

extern "C" int seqvp_capped_step(const double* limit, const double* root_out, double safety, double* tau,
                                 void* stream) {
  if (!limit || !root_out || !tau) return SEQVP_EINVAL;
  hipLaunchKernelGGL(capped_step_kernel, dim3(1), dim3(1), 0, (hipStream_t)stream, limit, root_out, safety, tau);
  return launch_status();
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
  hipLaunchKernelGGL(axpy_inject_kernel, dim3((nn + VB - 1) / VB), dim3(VB), 0, (hipStream_t)stream, theta, p, tau,
                     (const double*)nullptr, nx, ny, nz, theta_out, u);
  return launch_status();
}

extern "C" int seqvp_axpy_inject_dev(const double* theta, const double* p, const double* tau, int64_t nx, int64_t ny,
                                     int64_t nz, double* theta_out, double* u, void* stream) {
  if (!theta || !p || !tau || !u || nx < 3 || ny < 3 || nz < 3) return SEQVP_EINVAL;
  const int64_t nn = nx * ny * nz;
  hipLaunchKernelGGL(axpy_inject_kernel, dim3((nn + VB - 1) / VB), dim3(VB), 0, (hipStream_t)stream, theta, p, 0.0, tau,
                     nx, ny, nz, theta_out, u);
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

extern "C" int seqvp_bb2_products(const double* theta, const double* theta_old, const double* g, const double* g_old,
                                  const double* p_old, const double* p, int64_t n, double* out, void* ws,
                                  size_t ws_bytes, void* stream) {
  if (!theta || !theta_old || !g || !g_old || !p_old || !p || !out || n < 0) return SEQVP_EINVAL;
  if (!ws || ws_bytes < seqvp_vec_workspace_bytes(n)) return SEQVP_EWORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  double* part = (double*)ws;
  hipLaunchKernelGGL(bb2_kernel, dim3(VGRID), dim3(VB), 0, s, theta, theta_old, g, g_old, p_old, p, n, part);
  hipLaunchKernelGGL(dots_reduce_kernel, dim3(1), dim3(VB), 0, s, part, VGRID, 4, out);
  return launch_status();
}

extern "C" int seqvp_step_bound(const double* p, const double* u, double ox, double oy, double oz, double hx,
                                double hy, double hz, int64_t nx, int64_t ny, int64_t nz, double* out, void* ws,
                                size_t ws_bytes, void* stream) {
  if (!p || !u || !out || nx < 3 || ny < 3 || nz < 3) return SEQVP_EINVAL;
  if (!ws || ws_bytes < seqvp_vec_workspace_bytes(0)) return SEQVP_EWORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  double* part = (double*)ws;
  hipLaunchKernelGGL(step_bound_kernel, dim3(VGRID), dim3(VB), 0, s, p, u, ox, oy, oz, hx, hy, hz, nx, ny, nz, part);
  hipLaunchKernelGGL(fold_min_max_kernel, dim3(1), dim3(64), 0, s, part, VGRID, 0, out);
  return launch_status();
}

extern "C" int seqvp_max_abs_diff(const double* a, const double* b, int64_t n, double* out, void* ws, size_t ws_bytes,
                                  void* stream) {
  if (!a || !b || !out || n < 0) return SEQVP_EINVAL;
  if (!ws || ws_bytes < seqvp_vec_workspace_bytes(n)) return SEQVP_EWORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  double* part = (double*)ws;
  hipLaunchKernelGGL(maxabs_kernel, dim3(VGRID), dim3(VB), 0, s, a, b, n, part);
  hipLaunchKernelGGL(fold_min_max_kernel, dim3(1), dim3(64), 0, s, part, VGRID, 1, out);
  return launch_status();
}
