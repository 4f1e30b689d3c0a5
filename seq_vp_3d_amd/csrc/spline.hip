// Tensor-product cubic B-spline sampler (value + physical gradient) for gfx950.
//
// Evaluates the not-a-knot spline whose coefficients and knot vectors the host built once per
// moving image (interpolation.py:29-40) at node positions X + u (the warp) or at arbitrary points,
// as TorchSpline.sample does (jd_curl/three_d/backends/spline.py:12-34, 81-127): the span is
// searchsorted(t, x, right=True) - 1 clamped to [3, n-1], the four nonzero cubic basis values come
// from the Cox-de Boor triangle, their derivatives from the quadratic row, and 64 coefficient taps
// are contracted. One query per lane; ~1.1 kflop per point; coefficient taps are L2-resident
// gathers (neighbouring lanes hit neighbouring spans).
#include "q1_common.h"
#include "abi_util.h"

namespace seqvp {

// Cubic basis at x on knot vector t (nt knots, n = nt - 4 coefficients): first coefficient index,
// basis values b[0..3], derivatives db[0..3].
__device__ __forceinline__ int64_t cubic_basis(const double* __restrict__ t, int64_t nt, double x, double b[4],
                                               double db[4]) {
  const int64_t n = nt - 4;
  // searchsorted(t, x, right=True) - 1: the last index i with t[i] <= x (-1 when there is none, NaN included).
  // The knots are uniform away from the repeated ends, so a proportional guess lands on or next to the span; the
  // two walks then end with t[span] <= x < t[span + 1], which for sorted knots is exactly that index (the same
  // result as a bisection, with one or two independent loads instead of a chain of eight dependent ones).
  const double t0 = t[0], tl = t[nt - 1];
  int64_t span;
  if (!(x >= t0)) {
    span = -1;
  } else if (x >= tl) {
    span = nt - 1;
  } else {
    const double f = (x - t0) / (tl - t0) * (double)(nt - 1);
    span = f < 0.0 ? 0 : (f > (double)(nt - 2) ? nt - 2 : (int64_t)f);
    while (t[span + 1] <= x) ++span;   // ends: t[nt - 1] > x
    while (t[span] > x) --span;        // ends: t[0] <= x
  }
  span = span < 3 ? 3 : (span > n - 1 ? n - 1 : span);
  double basis[4] = {1.0, 0.0, 0.0, 0.0};
  double quad[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int degree = 1; degree <= 3; ++degree) {
    double values[4];
    double saved = 0.0;
#pragma unroll
    for (int r = 0; r < degree; ++r) {
      const double right = t[span + r + 1] - x;
      const double left = x - t[span + 1 - degree + r];
      const double temp = basis[r] / (right + left);
      values[r] = saved + right * temp;
      saved = left * temp;
    }
    values[degree] = saved;
#pragma unroll
    for (int r = 0; r <= degree; ++r) basis[r] = values[r];
    if (degree == 2) {
      quad[0] = values[0];
      quad[1] = values[1];
      quad[2] = values[2];
    }
  }
  const int64_t first = span - 3;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    b[i] = basis[i];
    const int64_t idx = first + i;
    const double a = t[idx + 3] - t[idx];
    const double bb = t[idx + 4] - t[idx + 1];
    const double left = i == 0 ? 0.0 : quad[i - 1];
    const double right = i == 3 ? 0.0 : quad[i];
    db[i] = 3.0 * left / (a != 0.0 ? a : 1.0) - 3.0 * right / (bb != 0.0 ? bb : 1.0);
  }
  return first;
}

struct SplineArgs {
  const double* coeff;
  int64_t cz, cy, cx;
  const double* tz;
  int64_t ntz;
  const double* ty;
  int64_t nty;
  const double* tx;
  int64_t ntx;
};

__device__ __forceinline__ void spline_eval(const SplineArgs& s, double px, double py, double pz, double* value,
                                            double grad[3], bool want_grad) {
  double bz[4], dz[4], by[4], dy[4], bx[4], dx[4];
  const int64_t iz = cubic_basis(s.tz, s.ntz, pz, bz, dz);
  const int64_t iy = cubic_basis(s.ty, s.nty, py, by, dy);
  const int64_t ix = cubic_basis(s.tx, s.ntx, px, bx, dx);
  double v = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    double v1 = 0.0, gx1 = 0.0, gy1 = 0.0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const double* row = s.coeff + ((iz + a) * s.cy + (iy + b)) * s.cx + ix;
      const double c0 = row[0], c1 = row[1], c2 = row[2], c3 = row[3];
      const double rv = (c0 * bx[0] + c1 * bx[1]) + (c2 * bx[2] + c3 * bx[3]);
      v1 += by[b] * rv;
      if (want_grad) {
        const double rd = (c0 * dx[0] + c1 * dx[1]) + (c2 * dx[2] + c3 * dx[3]);
        gx1 += by[b] * rd;
        gy1 += dy[b] * rv;
      }
    }
    v += bz[a] * v1;
    if (want_grad) {
      gx += bz[a] * gx1;
      gy += bz[a] * gy1;
      gz += dz[a] * v1;
    }
  }
  *value = v;
  if (want_grad) {
    grad[0] = gx;
    grad[1] = gy;
    grad[2] = gz;
  }
}

__device__ __forceinline__ void block_count(int bad, unsigned long long* status) {
  __shared__ int sh[4];
  int t = bad;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) t += __shfl_xor(t, m, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int tot = sh[0] + sh[1] + sh[2] + sh[3];
    if (tot) atomicAdd(status, (unsigned long long)tot);
  }
}

// Warp: queries at X + u on the node grid; free (interior) nodes need the open support, boundary
// nodes the closed one (interpolation.py:43-69, 109-113; residuals.py:222-225).
__global__ void __launch_bounds__(256) spline_warp_kernel(const double* __restrict__ u, SplineArgs s, double ox,
                                                          double oy, double oz, double hx, double hy, double hz,
                                                          int64_t nx, int64_t ny, int64_t nz,
                                                          double lox, double hix, double loy, double hiy, double loz,
                                                          double hiz, double* __restrict__ value,
                                                          double* __restrict__ grad,
                                                          unsigned long long* __restrict__ status) {
  const int64_t n = nx * ny * nz;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  int bad = 0;
  if (i < n) {
    const int64_t x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
    const double px = (ox + hx * (double)x) + u[3 * i];
    const double py = (oy + hy * (double)y) + u[3 * i + 1];
    const double pz = (oz + hz * (double)z) + u[3 * i + 2];
    const bool strict = x > 0 && x < nx - 1 && y > 0 && y < ny - 1 && z > 0 && z < nz - 1;
    bool ok = px >= lox && px <= hix && py >= loy && py <= hiy && pz >= loz && pz <= hiz;
    if (strict) ok = ok && px > lox && px < hix && py > loy && py < hiy && pz > loz && pz < hiz;
    bad = ok ? 0 : 1;
    double v = __builtin_nan(""), g[3] = {__builtin_nan(""), __builtin_nan(""), __builtin_nan("")};
    if (ok) spline_eval(s, px, py, pz, &v, g, grad != nullptr);
    if (value) value[i] = v;
    if (grad) {
      grad[3 * i] = g[0];
      grad[3 * i + 1] = g[1];
      grad[3 * i + 2] = g[2];
    }
  }
  block_count(bad, status);
}

__global__ void __launch_bounds__(256) spline_points_kernel(const double* __restrict__ pts, int64_t npts,
                                                            const uint8_t* __restrict__ strict, SplineArgs s,
                                                            double lox, double hix, double loy, double hiy,
                                                            double loz, double hiz, double* __restrict__ value,
                                                            double* __restrict__ grad,
                                                            unsigned long long* __restrict__ status) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  int bad = 0;
  if (i < npts) {
    const double px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
    bool ok = px >= lox && px <= hix && py >= loy && py <= hiy && pz >= loz && pz <= hiz;
    // a strict query must lie in the open support (TorchSpline.sample's strict mask, backends/spline.py:81-96)
    if (strict && strict[i]) ok = ok && px > lox && px < hix && py > loy && py < hiy && pz > loz && pz < hiz;
    bad = ok ? 0 : 1;
    double v = __builtin_nan(""), g[3] = {__builtin_nan(""), __builtin_nan(""), __builtin_nan("")};
    if (ok) spline_eval(s, px, py, pz, &v, g, grad != nullptr);
    if (value) value[i] = v;
    if (grad) {
      grad[3 * i] = g[0];
      grad[3 * i + 1] = g[1];
      grad[3 * i + 2] = g[2];
    }
  }
  block_count(bad, status);
}

}  // namespace seqvp

using namespace seqvp;

static bool spline_ok(const double* coeff, int64_t cz, int64_t cy, int64_t cx, const double* tz, int64_t ntz,
                      const double* ty, int64_t nty, const double* tx, int64_t ntx) {
  return coeff && tz && ty && tx && cz >= 4 && cy >= 4 && cx >= 4 && ntz == cz + 4 && nty == cy + 4 && ntx == cx + 4;
}

// Warp: the moving image lives on the same grid as the nodes, so the closed support per axis is
// [o, o + h (n-1)] with n the coefficient count (interpolation.py:43-69).
extern "C" int seqvp_spline_warp(const double* u, const double* coeff, int64_t cz, int64_t cy, int64_t cx,
                                 const double* tz, int64_t ntz, const double* ty, int64_t nty, const double* tx,
                                 int64_t ntx, double ox, double oy, double oz, double hx, double hy, double hz,
                                 int64_t nx, int64_t ny, int64_t nz, double* value, double* grad, int64_t* status,
                                 void* stream) {
  if (!u || !status || !spline_ok(coeff, cz, cy, cx, tz, ntz, ty, nty, tx, ntx)) return SEQVP_EINVAL;
  if (nx < 2 || ny < 2 || nz < 2) return SEQVP_EINVAL;
  const double hix = ox + hx * (double)(cx - 1), hiy = oy + hy * (double)(cy - 1), hiz = oz + hz * (double)(cz - 1);
  SplineArgs a{coeff, cz, cy, cx, tz, ntz, ty, nty, tx, ntx};
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(copy_or_zero_kernel, dim3(1), dim3(256), 0, s, nullptr, (double*)status, 1);
  const int64_t n = nx * ny * nz;
  hipLaunchKernelGGL(spline_warp_kernel, dim3((n + 255) / 256), dim3(256), 0, s, u, a, ox, oy, oz, hx, hy, hz, nx,
                     ny, nz, ox, hix, oy, hiy, oz, hiz, value, grad, (unsigned long long*)status);
  return launch_status();
}

extern "C" int seqvp_spline_points(const double* points, int64_t npts, const uint8_t* strict, const double* coeff,
                                   int64_t cz, int64_t cy,
                                   int64_t cx, const double* tz, int64_t ntz, const double* ty, int64_t nty,
                                   const double* tx, int64_t ntx, double lox, double hix, double loy, double hiy,
                                   double loz, double hiz, double* value, double* grad, int64_t* status,
                                   void* stream) {
  if (!points || !status || npts < 0 || !spline_ok(coeff, cz, cy, cx, tz, ntz, ty, nty, tx, ntx)) return SEQVP_EINVAL;
  SplineArgs a{coeff, cz, cy, cx, tz, ntz, ty, nty, tx, ntx};
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(copy_or_zero_kernel, dim3(1), dim3(256), 0, s, nullptr, (double*)status, 1);
  if (npts == 0) return launch_status();
  hipLaunchKernelGGL(spline_points_kernel, dim3((npts + 255) / 256), dim3(256), 0, s, points, npts, strict, a, lox,
                     hix, loy, hiy, loz, hiz, value, grad, (unsigned long long*)status);
  return launch_status();
}
