// Pyramid pre-optimiser on the device: the pre-objective P(u), its analytic gradient and the smoothed Adam
// step with the per-iteration move cap (gfx950), in float32 (the pre-optimiser) or float64 (the error
// reference). Restates jd_curl/three_d/preopt.py:157-233 (_pre_energy_tensors, _run_level_default) and the
// analytic gradient of preopt_grad.py:1-20, 120-205:
//
//   P(u) = 1/2 sum_nodes (T(x + u) - F)^2 + mu/2 sum_cells [(J_c - 1)^2 + |curl_c|^2]
//        + kappa sum_cells relu(j0 - J_c)^2 + alpha/2 sum_interior |h_c Lap_h u_c|^2
//
// u in voxel units on the interior nodes, (3, D-2, H-2, W-2) component-major (x, y, z), zero on the six
// faces; T = trilinear interpolation of the moving image with border clamping (align_corners); J_c/curl_c
// from the cell-centre derivatives (mean of the four parallel edge differences, F_cd = delta_cd +
// (h_c/h_d) du_c/dx_d). One iteration is five launches with no host synchronisation (the cap is applied
// from a device reduction), so a level can be captured in a graph and replayed.
//
// Sums: per-lane Kahan (float) or plain (double) partials, block and grid sums in double-double; min J by
// a fixed-order tree. Every reduction is deterministic.
#include <algorithm>

#include "q1_common.h"
#include "abi_util.h"

namespace seqvp {
namespace pre {

static constexpr int PB = 256;
static constexpr int PGRID = 1024;     // grid-stride launches: PGRID blocks at most (partials are sized by it)

struct PDims {
  int64_t D, H, W;          // nodes
  int64_t Di, Hi, Wi;       // interior nodes
  int64_t n, nint, ncell;
};

static inline PDims make_pdims(int64_t nx, int64_t ny, int64_t nz) {
  PDims d;
  d.D = nz; d.H = ny; d.W = nx;
  d.Di = nz - 2; d.Hi = ny - 2; d.Wi = nx - 2;
  d.n = nx * ny * nz;
  d.nint = d.Di * d.Hi * d.Wi;
  d.ncell = (nx - 1) * (ny - 1) * (nz - 1);
  return d;
}

static inline unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>((n + PB - 1) / PB, PGRID); }

// u component c at node (z, y, x) of the full grid (0 on the faces).
template <typename T>
__device__ __forceinline__ T uval(const T* __restrict__ u, int c, int64_t z, int64_t y, int64_t x, const PDims& d) {
  if (z <= 0 || z >= d.D - 1 || y <= 0 || y >= d.H - 1 || x <= 0 || x >= d.W - 1) return (T)0;
  return u[c * d.nint + ((z - 1) * d.Hi + (y - 1)) * d.Wi + (x - 1)];
}

// Per-lane accumulator: Kahan in float, plain in double.
template <typename T>
struct Acc {
  T s = 0, c = 0;
  __device__ __forceinline__ void add(T x) {
    if constexpr (sizeof(T) == 4) {
      const T y = x - c;
      const T t = s + y;
      c = (t - s) - y;
      s = t;
    } else {
      s += x;
    }
  }
  __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};

// Block double-double sums of NS lane values and min of NM lane values -> part[(k*nblocks + block)*2 (+1)].
template <int NS, int NM>
__device__ __forceinline__ void block_partials(const double* s, const double* m, double* __restrict__ part,
                                               int64_t nblocks) {
  __shared__ DD sh[PB / 64][NS > 0 ? NS : 1];
  __shared__ double shm[PB / 64][NM > 0 ? NM : 1];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const DD r = dd_wave_sum(DD{s[k], 0.0});
    if (lane == 0) sh[w][k] = r;
  }
#pragma unroll
  for (int k = 0; k < NM; ++k) {
    const double r = wave_nan_min(m[k]);
    if (lane == 0) shm[w][k] = r;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      DD a = sh[0][k];
      for (int j = 1; j < PB / 64; ++j) dd_add(a, sh[j][k]);
      part[2 * (k * nblocks + blockIdx.x)] = a.hi;
      part[2 * (k * nblocks + blockIdx.x) + 1] = a.lo;
    }
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      double a = shm[0][k];
      for (int j = 1; j < PB / 64; ++j) a = nan_min(a, shm[j][k]);
      part[2 * (NS * nblocks) + k * nblocks + blockIdx.x] = a;
    }
  }
}

// ---- data term: residual, its energy and the data gradient at the interior nodes ----------------------
template <typename T>
__global__ void __launch_bounds__(PB) voxel_kernel(const T* __restrict__ u, const T* __restrict__ F,
                                                   const T* __restrict__ M, PDims d, T* __restrict__ gd,
                                                   double* __restrict__ part) {
  Acc<T> acc;
  const T Wm = (T)(d.W - 1), Hm = (T)(d.H - 1), Dm = (T)(d.D - 1);
  for (int64_t v = (int64_t)blockIdx.x * PB + threadIdx.x; v < d.n; v += (int64_t)gridDim.x * PB) {
    const int64_t x = v % d.W, y = (v / d.W) % d.H, z = v / (d.W * d.H);
    const T ix_u = (T)x + uval(u, 0, z, y, x, d);
    const T iy_u = (T)y + uval(u, 1, z, y, x, d);
    const T iz_u = (T)z + uval(u, 2, z, y, x, d);
    const bool cx = ix_u > (T)0 && ix_u < Wm, cy = iy_u > (T)0 && iy_u < Hm, cz = iz_u > (T)0 && iz_u < Dm;
    const T ix = ix_u < (T)0 ? (T)0 : (ix_u > Wm ? Wm : ix_u);
    const T iy = iy_u < (T)0 ? (T)0 : (iy_u > Hm ? Hm : iy_u);
    const T iz = iz_u < (T)0 ? (T)0 : (iz_u > Dm ? Dm : iz_u);
    const T fx = floor(ix), fy = floor(iy), fz = floor(iz);
    const int64_t x0 = (int64_t)fx, y0 = (int64_t)fy, z0 = (int64_t)fz;
    const T wx1 = ix - fx, wy1 = iy - fy, wz1 = iz - fz;
    const T wx0 = (fx + (T)1) - ix, wy0 = (fy + (T)1) - iy, wz0 = (fz + (T)1) - iz;
    const bool inx = x0 + 1 <= d.W - 1, iny = y0 + 1 <= d.H - 1, inz = z0 + 1 <= d.D - 1;
    const int64_t b = (z0 * d.H + y0) * d.W + x0, sz = d.H * d.W;
    const T c000 = M[b];
    const T c001 = inx ? M[b + 1] : (T)0;
    const T c010 = iny ? M[b + d.W] : (T)0;
    const T c011 = (iny && inx) ? M[b + d.W + 1] : (T)0;
    const T c100 = inz ? M[b + sz] : (T)0;
    const T c101 = (inz && inx) ? M[b + sz + 1] : (T)0;
    const T c110 = (inz && iny) ? M[b + sz + d.W] : (T)0;
    const T c111 = (inz && iny && inx) ? M[b + sz + d.W + 1] : (T)0;
    T val = c000 * ((wx0 * wy0) * wz0);
    val += c001 * ((wx1 * wy0) * wz0);
    val += c010 * ((wx0 * wy1) * wz0);
    val += c011 * ((wx1 * wy1) * wz0);
    val += c100 * ((wx0 * wy0) * wz1);
    val += c101 * ((wx1 * wy0) * wz1);
    val += c110 * ((wx0 * wy1) * wz1);
    val += c111 * ((wx1 * wy1) * wz1);
    const T r = val - F[v];
    acc.add(r * r);
    if (x > 0 && x < d.W - 1 && y > 0 && y < d.H - 1 && z > 0 && z < d.D - 1) {
      const T dtx = (c001 - c000) * (wy0 * wz0) + (c011 - c010) * (wy1 * wz0) + (c101 - c100) * (wy0 * wz1) +
                    (c111 - c110) * (wy1 * wz1);
      const T dty = (c010 - c000) * (wx0 * wz0) + (c011 - c001) * (wx1 * wz0) + (c110 - c100) * (wx0 * wz1) +
                    (c111 - c101) * (wx1 * wz1);
      const T dtz = (c100 - c000) * (wx0 * wy0) + (c101 - c001) * (wx1 * wy0) + (c110 - c010) * (wx0 * wy1) +
                    (c111 - c011) * (wx1 * wy1);
      const int64_t k = ((z - 1) * d.Hi + (y - 1)) * d.Wi + (x - 1);
      gd[k] = cx ? r * dtx : (T)0;
      gd[d.nint + k] = cy ? r * dty : (T)0;
      gd[2 * d.nint + k] = cz ? r * dtz : (T)0;
    }
  }
  const double s[1] = {acc.value()};
  block_partials<1, 0>(s, nullptr, part, gridDim.x);
}

// ---- geometry and floor terms: per-cell J, curl and G = dE_cell / d(cell derivative) -------------------
template <typename T>
__global__ void __launch_bounds__(PB) cell_kernel(const T* __restrict__ u, PDims d, T hx, T hy, T hz, T mu,
                                                  T kappa, T j0, T* __restrict__ G, double* __restrict__ part) {
  Acc<T> a_det, a_curl, a_floor;
  double mn = __builtin_inf();
  const int64_t cxn = d.W - 1, cyn = d.H - 1;
  const T h[3] = {hx, hy, hz};
  for (int64_t c = (int64_t)blockIdx.x * PB + threadIdx.x; c < d.ncell; c += (int64_t)gridDim.x * PB) {
    const int64_t x = c % cxn, y = (c / cxn) % cyn, z = c / (cxn * cyn);
    T der[3][3];   // der[comp][axis x,y,z] voxel derivatives at the cell centre
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const T a000 = uval(u, k, z, y, x, d), a001 = uval(u, k, z, y, x + 1, d);
      const T a010 = uval(u, k, z, y + 1, x, d), a011 = uval(u, k, z, y + 1, x + 1, d);
      const T a100 = uval(u, k, z + 1, y, x, d), a101 = uval(u, k, z + 1, y, x + 1, d);
      const T a110 = uval(u, k, z + 1, y + 1, x, d), a111 = uval(u, k, z + 1, y + 1, x + 1, d);
      // edge differences along the axis, averaged over z first, then over the remaining axis
      der[k][0] = (T)0.5 * ((T)0.5 * ((a001 - a000) + (a101 - a100)) + (T)0.5 * ((a011 - a010) + (a111 - a110)));
      der[k][1] = (T)0.5 * ((T)0.5 * ((a010 - a000) + (a110 - a100)) + (T)0.5 * ((a011 - a001) + (a111 - a101)));
      der[k][2] = (T)0.5 * ((T)0.5 * ((a100 - a000) + (a110 - a010)) + (T)0.5 * ((a101 - a001) + (a111 - a011)));
    }
    T R[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int j = 0; j < 3; ++j) R[k][j] = h[k] / h[j];
    const T F00 = (T)1 + der[0][0], F01 = der[0][1] * R[0][1], F02 = der[0][2] * R[0][2];
    const T F10 = der[1][0] * R[1][0], F11 = (T)1 + der[1][1], F12 = der[1][2] * R[1][2];
    const T F20 = der[2][0] * R[2][0], F21 = der[2][1] * R[2][1], F22 = (T)1 + der[2][2];
    const T C00 = F11 * F22 - F12 * F21, C01 = -(F10 * F22 - F12 * F20), C02 = F10 * F21 - F11 * F20;
    const T C10 = -(F01 * F22 - F02 * F21), C11 = F00 * F22 - F02 * F20, C12 = -(F00 * F21 - F01 * F20);
    const T C20 = F01 * F12 - F02 * F11, C21 = -(F00 * F12 - F02 * F10), C22 = F00 * F11 - F01 * F10;
    const T det = F00 * C00 + F01 * C01 + F02 * C02;
    const T k0 = F21 - F12, k1 = F02 - F20, k2 = F10 - F01;
    const T rl = j0 - det > (T)0 ? j0 - det : (T)0;
    const T A = mu * (det - (T)1) - (T)2 * kappa * rl;
    a_det.add((det - (T)1) * (det - (T)1));
    a_curl.add(k0 * k0 + k1 * k1 + k2 * k2);
    a_floor.add(rl * rl);
    mn = nan_min(mn, (double)det);
    G[0 * d.ncell + c] = A * C00;
    G[1 * d.ncell + c] = (A * C01 + mu * (-k2)) * R[0][1];
    G[2 * d.ncell + c] = (A * C02 + mu * k1) * R[0][2];
    G[3 * d.ncell + c] = (A * C10 + mu * k2) * R[1][0];
    G[4 * d.ncell + c] = A * C11;
    G[5 * d.ncell + c] = (A * C12 + mu * (-k0)) * R[1][2];
    G[6 * d.ncell + c] = (A * C20 + mu * (-k1)) * R[2][0];
    G[7 * d.ncell + c] = (A * C21 + mu * k0) * R[2][1];
    G[8 * d.ncell + c] = A * C22;
  }
  const double s[3] = {a_det.value(), a_curl.value(), a_floor.value()};
  const double m[1] = {mn};
  block_partials<3, 1>(s, m, part, gridDim.x);
}

// seven-point Laplacian (physical) of the interior nodes, times h_c: lap (3, nint)
template <typename T>
__device__ __forceinline__ T lap_at(const T* __restrict__ u, int c, int64_t z, int64_t y, int64_t x, const PDims& d,
                                    T hx, T hy, T hz) {
  const T ui = uval(u, c, z, y, x, d);
  return (uval(u, c, z + 1, y, x, d) - (T)2 * ui + uval(u, c, z - 1, y, x, d)) / (hz * hz) +
         (uval(u, c, z, y + 1, x, d) - (T)2 * ui + uval(u, c, z, y - 1, x, d)) / (hy * hy) +
         (uval(u, c, z, y, x + 1, d) - (T)2 * ui + uval(u, c, z, y, x - 1, d)) / (hx * hx);
}

template <typename T>
__global__ void __launch_bounds__(PB) lap_kernel(const T* __restrict__ u, PDims d, T hx, T hy, T hz,
                                                 T* __restrict__ lap, double* __restrict__ part) {
  Acc<T> acc;
  const T h[3] = {hx, hy, hz};
  for (int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x; i < d.nint; i += (int64_t)gridDim.x * PB) {
    const int64_t x = i % d.Wi + 1, y = (i / d.Wi) % d.Hi + 1, z = i / (d.Wi * d.Hi) + 1;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T l = lap_at(u, c, z, y, x, d, hx, hy, hz) * h[c];
      lap[c * d.nint + i] = l;
      acc.add(l * l);
    }
  }
  const double s[1] = {acc.value()};
  block_partials<1, 0>(s, nullptr, part, gridDim.x);
}

// grad (3, nint) = data gradient + sum of the 8 cells' G with +-1/4 per axis + alpha h_c Lap_h(lap_c)
// (in place over gd).
template <typename T>
__global__ void __launch_bounds__(PB) node_kernel(T* __restrict__ g, const T* __restrict__ G,
                                                  const T* __restrict__ lap, PDims d, T alpha, T hx, T hy, T hz) {
  const int64_t cxn = d.W - 1, cyn = d.H - 1;
  const T h[3] = {hx, hy, hz};
  for (int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x; i < d.nint; i += (int64_t)gridDim.x * PB) {
    const int64_t x = i % d.Wi + 1, y = (i / d.Wi) % d.Hi + 1, z = i / (d.Wi * d.Hi) + 1;
    T acc[3] = {g[i], g[d.nint + i], g[2 * d.nint + i]};
#pragma unroll
    for (int oz = 0; oz < 2; ++oz)
#pragma unroll
      for (int oy = 0; oy < 2; ++oy)
#pragma unroll
        for (int ox = 0; ox < 2; ++ox) {
          const int64_t c = ((z - oz) * cyn + (y - oy)) * cxn + (x - ox);
          const T sx = ox ? (T)0.25 : (T)-0.25, sy = oy ? (T)0.25 : (T)-0.25, sz = oz ? (T)0.25 : (T)-0.25;
#pragma unroll
          for (int k = 0; k < 3; ++k)
            acc[k] += sx * G[(3 * k + 0) * d.ncell + c] + sy * G[(3 * k + 1) * d.ncell + c] +
                      sz * G[(3 * k + 2) * d.ncell + c];
        }
    if (alpha > (T)0) {
      // Lap_h of lap with zero faces: lap is stored on the interior; reuse uval's face rule
#pragma unroll
      for (int k = 0; k < 3; ++k) acc[k] += alpha * h[k] * lap_at(lap, k, z, y, x, d, hx, hy, hz);
    }
    g[i] = acc[0];
    g[d.nint + i] = acc[1];
    g[2 * d.nint + i] = acc[2];
  }
}

// terms = [total, data, geometry, floor, curvature, min J] from the partials of the three sum kernels.
__global__ void __launch_bounds__(PB) terms_kernel(const double* __restrict__ pv, int64_t nv,
                                                   const double* __restrict__ pc, int64_t nc,
                                                   const double* __restrict__ pl, int64_t nl, double mu,
                                                   double kappa, double alpha, double* __restrict__ terms) {
  // sums: data (pv 0), det2/curl2/floor (pc 0..2), lap2 (pl 0); min J (pc after the sums)
  __shared__ DD sh[PB / 64][5];
  __shared__ double shm[PB / 64];
  DD s[5] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
  double m = __builtin_inf();
  for (int64_t i = threadIdx.x; i < nv; i += PB) dd_add(s[0], DD{pv[2 * i], pv[2 * i + 1]});
  for (int k = 0; k < 3; ++k)
    for (int64_t i = threadIdx.x; i < nc; i += PB)
      dd_add(s[1 + k], DD{pc[2 * (k * nc + i)], pc[2 * (k * nc + i) + 1]});
  for (int64_t i = threadIdx.x; i < nl; i += PB) dd_add(s[4], DD{pl[2 * i], pl[2 * i + 1]});
  for (int64_t i = threadIdx.x; i < nc; i += PB) m = nan_min(m, pc[2 * 3 * nc + i]);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int k = 0; k < 5; ++k) {
    DD r = s[k];
    for (int o = 32; o >= 1; o >>= 1) dd_add(r, DD{__shfl_xor(r.hi, o, 64), __shfl_xor(r.lo, o, 64)});
    if (lane == 0) sh[w][k] = r;
  }
  m = wave_nan_min(m);
  if (lane == 0) shm[w] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v[5];
    for (int k = 0; k < 5; ++k) {
      DD a = sh[0][k];
      for (int j = 1; j < PB / 64; ++j) dd_add(a, sh[j][k]);
      v[k] = a.hi + a.lo;
    }
    double mm = shm[0];
    for (int j = 1; j < PB / 64; ++j) mm = nan_min(mm, shm[j]);
    const double data = 0.5 * v[0];
    const double geometry = 0.5 * mu * (v[1] + v[2]);
    const double floor_ = kappa * v[3];
    const double curvature = alpha > 0.0 ? 0.5 * alpha * v[4] : 0.0;
    terms[0] = ((data + geometry) + floor_) + curvature;
    terms[1] = data;
    terms[2] = geometry;
    terms[3] = floor_;
    terms[4] = curvature;
    terms[5] = mm;
  }
}

// ---- the update: Gaussian smoothing, Adam, cap -------------------------------------------------------------
// out = separable smoothing of in along `axis` (0 = z, 1 = y, 2 = x) of a (3, Di, Hi, Wi) array, replicate
// padding, weights w[0..2r].
template <typename T>
__global__ void __launch_bounds__(PB) smooth_kernel(const T* __restrict__ in, T* __restrict__ out, PDims d, int axis,
                                                    const T* __restrict__ w, int r) {
  const int64_t total = 3 * d.nint;
  const int64_t len = axis == 0 ? d.Di : (axis == 1 ? d.Hi : d.Wi);
  const int64_t stride = axis == 0 ? d.Hi * d.Wi : (axis == 1 ? d.Wi : 1);
  for (int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x; i < total; i += (int64_t)gridDim.x * PB) {
    const int64_t pos = (i / stride) % len;
    const int64_t base = i - pos * stride;
    T acc = (T)0;
    for (int k = -r; k <= r; ++k) {
      int64_t q = pos + k;
      q = q < 0 ? 0 : (q >= len ? len - 1 : q);
      acc += w[k + r] * in[base + q * stride];
    }
    out[i] = acc;
  }
}

// m, v <- Adam moments of g; step = lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps) with t = state[0] + 1;
// block maxima of |step| -> pmax.
template <typename T>
__global__ void __launch_bounds__(PB) adam_kernel(const T* __restrict__ g, T* __restrict__ m, T* __restrict__ v,
                                                  T* __restrict__ step, int64_t n, T lr, T b1, T b2, T eps,
                                                  const long long* __restrict__ state, T* __restrict__ pmax) {
  const double t = (double)(state[0] + 1);
  const T c1 = (T)(1.0 - pow((double)b1, t)), c2 = (T)(1.0 - pow((double)b2, t));
  T best = (T)0;
  for (int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x; i < n; i += (int64_t)gridDim.x * PB) {
    const T gi = g[i];
    const T mi = b1 * m[i] + ((T)1 - b1) * gi;
    const T vi = b2 * v[i] + ((T)1 - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const T s = lr * (mi / c1) / (sqrt(vi / c2) + eps);
    step[i] = s;
    const T a = fabs(s);
    best = (a > best || a != a) ? a : best;
  }
  for (int o = 32; o >= 1; o >>= 1) {
    const T other = __shfl_xor(best, o, 64);
    best = (other > best || other != other) ? other : best;
  }
  __shared__ T sh[PB / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int j = 1; j < PB / 64; ++j) best = (sh[j] > best || sh[j] != sh[j]) ? sh[j] : best;
    pmax[blockIdx.x] = best;
  }
}

// u -= step * scale, scale = 1 if max|step| <= cap else cap / max|step|; state[0] += 1 (t), state[1] += (scale < 1).
template <typename T>
__global__ void __launch_bounds__(PB) apply_kernel(T* __restrict__ u, const T* __restrict__ step, int64_t n,
                                                   const T* __restrict__ pmax, int nmax, T cap,
                                                   long long* __restrict__ state) {
  __shared__ T sh[PB / 64];
  T best = (T)0;
  for (int i = threadIdx.x; i < nmax; i += PB) best = (pmax[i] > best || pmax[i] != pmax[i]) ? pmax[i] : best;
  for (int o = 32; o >= 1; o >>= 1) {
    const T other = __shfl_xor(best, o, 64);
    best = (other > best || other != other) ? other : best;
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = best;
  __syncthreads();
  T largest = sh[0];
  for (int j = 1; j < PB / 64; ++j) largest = (sh[j] > largest || sh[j] != sh[j]) ? sh[j] : largest;
  const T scale = largest <= cap ? (T)1 : (T)((double)cap / (double)largest);
  for (int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x; i < n; i += (int64_t)gridDim.x * PB) u[i] -= step[i] * scale;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    state[0] += 1;
    if (scale < (T)1) state[1] += 1;
  }
}

template <typename T>
__global__ void __launch_bounds__(PB) copy_kernel(const T* __restrict__ in, T* __restrict__ out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x; i < n; i += (int64_t)gridDim.x * PB) out[i] = in[i];
}

// workspace layout (bytes): G (9 ncell T) | lap (3 nint T) | partials (3 regions of 8*PGRID doubles)
static inline size_t ws_bytes(const PDims& d, size_t tsize) {
  const size_t a = (9 * (size_t)d.ncell + 3 * (size_t)d.nint) * tsize;
  return ((a + 255) / 256) * 256 + 3 * 8 * PGRID * sizeof(double);
}

template <typename T>
static int energy_grad(const T* u, const T* F, const T* M, const PDims& d, double hx, double hy, double hz, double mu,
                       double alpha, double kappa, double j0, T* grad, double* terms, void* ws, hipStream_t s) {
  const size_t a = (9 * (size_t)d.ncell + 3 * (size_t)d.nint) * sizeof(T);
  T* G = (T*)ws;
  T* lap = G + 9 * d.ncell;
  double* pv = (double*)((char*)ws + ((a + 255) / 256) * 256);
  double* pc = pv + 8 * PGRID;
  double* pl = pc + 8 * PGRID;
  const unsigned gv = grid_for(d.n), gc = grid_for(d.ncell), gl = grid_for(d.nint);
  hipLaunchKernelGGL(voxel_kernel<T>, dim3(gv), dim3(PB), 0, s, u, F, M, d, grad, pv);
  hipLaunchKernelGGL(cell_kernel<T>, dim3(gc), dim3(PB), 0, s, u, d, (T)hx, (T)hy, (T)hz, (T)mu, (T)kappa, (T)j0, G,
                     pc);
  hipLaunchKernelGGL(lap_kernel<T>, dim3(gl), dim3(PB), 0, s, u, d, (T)hx, (T)hy, (T)hz, lap, pl);
  hipLaunchKernelGGL(node_kernel<T>, dim3(gl), dim3(PB), 0, s, grad, G, lap, d, (T)alpha, (T)hx, (T)hy, (T)hz);
  hipLaunchKernelGGL(terms_kernel, dim3(1), dim3(PB), 0, s, pv, (int64_t)gv, pc, (int64_t)gc, pl, (int64_t)gl, mu,
                     kappa, alpha, terms);
  return launch_status();
}

}  // namespace pre
}  // namespace seqvp

using namespace seqvp::pre;

extern "C" size_t seqvp_preopt_workspace_bytes(int64_t nx, int64_t ny, int64_t nz, int fp32) {
  if (nx < 3 || ny < 3 || nz < 3) return 0;
  const PDims d = make_pdims(nx, ny, nz);
  return ws_bytes(d, fp32 ? sizeof(float) : sizeof(double));
}

extern "C" int seqvp_preopt_energy_grad(const void* u, const void* fixed, const void* moving, int64_t nx, int64_t ny,
                                        int64_t nz, double hx, double hy, double hz, double mu, double alpha,
                                        double kappa, double j0, int fp32, void* grad, double* terms, void* workspace,
                                        size_t workspace_bytes, void* stream) {
  if (!u || !fixed || !moving || !grad || !terms || nx < 3 || ny < 3 || nz < 3) return SEQVP_EINVAL;
  if (!(hx > 0 && hy > 0 && hz > 0)) return SEQVP_EINVAL;
  const PDims d = make_pdims(nx, ny, nz);
  if (!workspace || workspace_bytes < ws_bytes(d, fp32 ? sizeof(float) : sizeof(double))) return SEQVP_EWORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  if (fp32)
    return energy_grad<float>((const float*)u, (const float*)fixed, (const float*)moving, d, hx, hy, hz, mu, alpha,
                              kappa, j0, (float*)grad, terms, workspace, s);
  return energy_grad<double>((const double*)u, (const double*)fixed, (const double*)moving, d, hx, hy, hz, mu, alpha,
                             kappa, j0, (double*)grad, terms, workspace, s);
}

template <typename T>
static int preopt_step(T* u, T* m, T* v, T* g, T* tmp, const PDims& d, const T* wg, int rg, const T* wf, int rf,
                       double lr, double cap, double b1, double b2, double eps, long long* state, T* pmax,
                       hipStream_t s) {
  const int64_t n = 3 * d.nint;
  const unsigned gn = grid_for(n);
  // smoothed gradient: g -> tmp -> g -> tmp (three axes); no smoothing: tmp = g
  const T* gs = g;
  if (rg >= 0) {
    hipLaunchKernelGGL(smooth_kernel<T>, dim3(gn), dim3(PB), 0, s, g, tmp, d, 0, wg, rg);
    hipLaunchKernelGGL(smooth_kernel<T>, dim3(gn), dim3(PB), 0, s, tmp, g, d, 1, wg, rg);
    hipLaunchKernelGGL(smooth_kernel<T>, dim3(gn), dim3(PB), 0, s, g, tmp, d, 2, wg, rg);
    gs = tmp;
  }
  // the step overwrites the gradient buffer not holding gs
  T* step = gs == g ? tmp : g;
  hipLaunchKernelGGL(adam_kernel<T>, dim3(gn), dim3(PB), 0, s, gs, m, v, step, n, (T)lr, (T)b1, (T)b2, (T)eps,
                     (const long long*)state, pmax);
  hipLaunchKernelGGL(apply_kernel<T>, dim3(gn), dim3(PB), 0, s, u, step, n, (const T*)pmax, (int)gn, (T)cap, state);
  if (rf >= 0) {
    T* other = step;    // free now
    hipLaunchKernelGGL(smooth_kernel<T>, dim3(gn), dim3(PB), 0, s, u, other, d, 0, wf, rf);
    T* other2 = other == g ? tmp : g;
    hipLaunchKernelGGL(smooth_kernel<T>, dim3(gn), dim3(PB), 0, s, other, other2, d, 1, wf, rf);
    hipLaunchKernelGGL(smooth_kernel<T>, dim3(gn), dim3(PB), 0, s, other2, u, d, 2, wf, rf);
  }
  return launch_status();
}

extern "C" int seqvp_preopt_step(void* u, void* m, void* v, void* grad, void* tmp, int64_t nx, int64_t ny, int64_t nz,
                                 const void* w_grad, int r_grad, const void* w_field, int r_field, double lr,
                                 double cap, double beta1, double beta2, double eps, int fp32, long long* state,
                                 void* pmax, void* stream) {
  if (!u || !m || !v || !grad || !tmp || !state || !pmax || nx < 3 || ny < 3 || nz < 3) return SEQVP_EINVAL;
  if ((r_grad >= 0 && !w_grad) || (r_field >= 0 && !w_field) || r_grad > 64 || r_field > 64) return SEQVP_EINVAL;
  const PDims d = make_pdims(nx, ny, nz);
  hipStream_t s = (hipStream_t)stream;
  if (fp32)
    return preopt_step<float>((float*)u, (float*)m, (float*)v, (float*)grad, (float*)tmp, d, (const float*)w_grad,
                              r_grad, (const float*)w_field, r_field, lr, cap, beta1, beta2, eps, state, (float*)pmax,
                              s);
  return preopt_step<double>((double*)u, (double*)m, (double*)v, (double*)grad, (double*)tmp, d,
                             (const double*)w_grad, r_grad, (const double*)w_field, r_field, lr, cap, beta1, beta2,
                             eps, state, (double*)pmax, s);
}
