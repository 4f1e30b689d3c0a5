// DST-I Poisson inverse on MFMA (gfx950).
//
// p = sign * s * IDST( DST(s * g) / den ) with DST the orthonormal type-I transform along x, y and z
// (poisson.py:99-106; torch_ops.py:98-115; hybrid.py:262-286). The interior lengths of the native
// grid are 222/190/158 (periods 223, 191 prime, 159 = 3*53), so there is no small-radix FFT; each
// axis is a dense contraction with the symmetric sine matrix S[k][j] = sqrt(2/(n+1)) sin(pi (j+1)(k+1)/(n+1)),
// run as batched GEMMs on the matrix cores:
//   x axis (contiguous):  Out (O x n) = X (O x n) * S            (O = ncomp*mz*my rows)
//   y/z axes (strided):   Out_o (n x I) = S * X_o (n x I)         (I = inner stride, batch o)
// The pre-scale s*g is fused into the first pass's loads, the division by the spectrum into the
// third pass's stores and the post-scale/negation into the last pass's stores: 6 passes over the
// 3F-vector, ping-ponging through one workspace buffer.
//
// Tiles: lengths up to 256 use the folded strip kernels below (a workgroup owns 64 lines of the independent index and
// one wave per 16 outputs, v_mfma_f64_16x16x4f64); longer axes use a 64x64 output tile per workgroup of 4 wave64,
// each wave a 32x32 quadrant as 2x2 accumulators. Operand fragments of the data matrix are staged through LDS in
// 64x16 (or 16x64) k-slabs with coalesced global loads; S fragments are read from L2.
#include <math.h>

#include "q1_common.h"
#include "abi_util.h"

namespace seqvp {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int TM = 64, TN = 64, TK = 16;

template <typename T> struct Acc;
template <> struct Acc<double> { typedef d4 type; };
template <> struct Acc<float> { typedef f4 type; };

__device__ __forceinline__ d4 mfma(double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ __forceinline__ f4 mfma(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// Operands (both types): lane l holds A[l%16][l/16] and B[l/16][l%16]. Accumulator row of register
// r in lane group lg = l/16 (column l%16): interleaved lg + 4r for the f64 instruction, blocked
// 4lg + r for the f32 one (checked on the device by seqvp_mfma_probe).
template <typename T> __device__ __forceinline__ int acc_row(int lg, int r);
template <> __device__ __forceinline__ int acc_row<double>(int lg, int r) { return lg + 4 * r; }
template <> __device__ __forceinline__ int acc_row<float>(int lg, int r) { return 4 * lg + r; }

// sine matrix of length n into S (ld x ld, ld >= n, zero outside n x n), row-major, exact angle reduction
// m = (j+1)(k+1) mod 2(n+1), evaluated in binary64 and rounded once to T
template <typename T>
__global__ void sine_matrix_kernel(T* __restrict__ S, int64_t n, int64_t ld) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= ld * ld) return;
  const int64_t k = i / ld, j = i % ld;
  if (k >= n || j >= n) {
    S[i] = (T)0;
    return;
  }
  const int64_t period = 2 * (n + 1);
  const int64_t m = ((j + 1) * (k + 1)) % period;
  S[i] = (T)(sqrt(2.0 / (double)(n + 1)) * sinpi((double)m / (double)(n + 1)));
}

// Epilogue/prologue modes.
enum : int { PRE_NONE = 0, PRE_SCALE = 1 };
enum : int { POST_NONE = 0, POST_DIV = 1, POST_SCALE = 2 };

// ---- x axis: Out[o][kk] = sum_j X[o][j] S[j][kk]  (M = O rows, N = K = n) --------------------------
// A = X (row-major, ld n), B = S. Block tile: rows m0..m0+63, cols n0..n0+63. PRE_SCALE forms
// s*x in binary64 before rounding to T (the equilibration pre-scale); POST_SCALE writes
// sign*(s*y) in binary64.
template <typename T, typename TIN, typename TOUT, int PRE, int POST>
__global__ void __launch_bounds__(256) dst_rows_kernel(const TIN* __restrict__ X, const double* __restrict__ pre,
                                                       const T* __restrict__ S, int64_t O, int64_t n,
                                                       const double* __restrict__ post, double sign,
                                                       TOUT* __restrict__ Y) {
  static_assert(POST != POST_DIV, "the spectral division is fused into a strided pass");
  typedef typename Acc<T>::type AT;
  __shared__ T As[TK][TM + 1];   // As[k][m]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // 1-D grid over (row tile, column tile), column tile fastest; each XCD takes a contiguous range of tiles so the
  // column tiles that re-read the same X rows run back to back on one L2.
  const int64_t ncb = (n + TN - 1) / TN;
  const int64_t tile = xcd_block(blockIdx.x, gridDim.x);
  const int64_t m0 = (tile / ncb) * TM, n0 = (tile % ncb) * TN;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  AT acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = AT{0, 0, 0, 0};
  const int li = lane & 15, lk = lane >> 4;
  for (int64_t k0 = 0; k0 < n; k0 += TK) {
    // stage A tile: 64 rows x 16 k, coalesced along k (row-major X)
#pragma unroll
    for (int t = 0; t < (TM * TK) / 256; ++t) {
      const int e = t * 256 + threadIdx.x;
      const int r = e / TK, kk = e % TK;
      const int64_t gm = m0 + r, gk = k0 + kk;
      T v = 0;
      if (gm < O && gk < n) {
        const int64_t idx = gm * n + gk;
        if (PRE == PRE_SCALE)
          v = (T)(pre[idx] * (double)X[idx]);
        else
          v = (T)X[idx];
      }
      As[kk][r] = v;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < TK; ks += 4) {
      const int64_t gk = k0 + ks + lk;
      T bf[2];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int64_t gn = n0 + wn + 16 * b + li;
        bf[b] = (gk < n && gn < n) ? S[gk * n + gn] : (T)0;
      }
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const T af = As[ks + lk][wm + 16 * a + li];
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = mfma(af, bf[b], acc[a][b]);
      }
    }
    __syncthreads();
  }
  // accumulator layout: lane holds column li, row acc_row<T>(lk, r)
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gm = m0 + wm + 16 * a + acc_row<T>(lk, r), gn = n0 + wn + 16 * b + li;
        if (gm < O && gn < n) {
          const int64_t idx = gm * n + gn;
          if (POST == POST_SCALE) {
            const double v = (double)acc[a][b][r];
            Y[idx] = (TOUT)(sign * (post ? post[idx] * v : v));
          } else {
            Y[idx] = (TOUT)acc[a][b][r];
          }
        }
      }
}

// ---- strided axes: Out_o[kk][i] = sum_j S[kk][j] X_o[j][i]  (M = n, N = I, K = n, batch o) ----------
// A = S (L2), B = X_o staged in LDS as Bs[k][i] (coalesced along i). With POST_DIV the result is
// divided by den[kk*I + i] rounded to T (the z pass: den has the interior shape).
template <typename T, int POST>
__global__ void __launch_bounds__(256) dst_cols_kernel(const T* __restrict__ X, const T* __restrict__ S, int64_t n,
                                                       int64_t I, const double* __restrict__ den,
                                                       T* __restrict__ Y) {
  typedef typename Acc<T>::type AT;
  __shared__ T Bs[TK][TN + 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // 1-D grid over (batch, inner tile, output-row tile), output-row tile fastest, XCD-contiguous (the output-row
  // tiles of one inner tile re-read the same X columns).
  const int64_t nmb = (n + TM - 1) / TM, nib = (I + TN - 1) / TN;
  const int64_t tile = xcd_block(blockIdx.x, gridDim.x);
  const int64_t o = tile / (nmb * nib);
  const int64_t rest = tile % (nmb * nib);
  const int64_t m0 = (rest % nmb) * TM, n0 = (rest / nmb) * TN;
  const T* Xo = X + o * n * I;
  T* Yo = Y + o * n * I;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  AT acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = AT{0, 0, 0, 0};
  const int li = lane & 15, lk = lane >> 4;
  for (int64_t k0 = 0; k0 < n; k0 += TK) {
#pragma unroll
    for (int t = 0; t < (TK * TN) / 256; ++t) {
      const int e = t * 256 + threadIdx.x;
      const int kk = e / TN, c = e % TN;
      const int64_t gk = k0 + kk, gi = n0 + c;
      Bs[kk][c] = (gk < n && gi < I) ? Xo[gk * I + gi] : (T)0;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < TK; ks += 4) {
      const int64_t gk = k0 + ks + lk;
      T af[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const int64_t gm = m0 + wm + 16 * a + li;
        af[a] = (gm < n && gk < n) ? S[gm * n + gk] : (T)0;
      }
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const T bf = Bs[ks + lk][wn + 16 * b + li];
#pragma unroll
        for (int a = 0; a < 2; ++a) acc[a][b] = mfma(af[a], bf, acc[a][b]);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gm = m0 + wm + 16 * a + acc_row<T>(lk, r), gi = n0 + wn + 16 * b + li;
        if (gm < n && gi < I) {
          const int64_t idx = gm * I + gi;
          T v = acc[a][b][r];
          if (POST == POST_DIV) v = v / (T)den[idx];
          Yo[idx] = v;
        }
      }
}

// ---- folded passes (n <= 256) ------------------------------------------------------------------------------------
// The DST-I matrix satisfies S[n-1-j][k] = (-1)^k S[j][k]. Pairing j with n-1-j, an output k is
//   even k: sum_{j < h'} U[j] S[j][k],  U[j] = x[j] + x[n-1-j] (j < h), U[h] = x[h] when n is odd,
//   odd k:  sum_{j < h}  V[j] S[j][k],  V[j] = x[j] - x[n-1-j],
// with h = floor(n/2), h' = ceil(n/2): two half-length contractions, half the MFMA work of the plain one.
// U and V are formed while a k-slab (FK = 16 folded indices) is written to LDS, double-buffered: the next slab's
// global loads are in flight in registers while the waves run the current slab's MFMAs, one barrier per slab. Every
// load is unconditional from a clamped in-bounds address (zero padding by select), the sine fragments of a slab are
// issued before its first MFMA, and the sine matrix is stored zero-padded, so no wait sits between a load and the
// MFMAs it overlaps.
constexpr int FK = 16;

__device__ __forceinline__ int64_t clamp_index(int64_t i, int64_t n) { return i < n ? i : n - 1; }

// Strip tiling: the outputs are cut into 16-wide strips: ne = ceil(ceil(n/2)/16) strips of even
// outputs, then no = ceil(floor(n/2)/16) strips of odd outputs. One workgroup of NW = max(ne + no, 2) waves owns a
// 64-wide tile of the independent index, and wave w computes strip w against the whole tile (four v_mfma
// accumulators), so the MFMA work is that of ceil(n/2) outputs rounded up to 16 per side (lengths 222/190/158:
// 224/192/160 output rows; a 64-output tiling per wave would compute 256). Per apply at 160x192x224: 0.92 ms against
// 0.95 ms for the 64-output tiling (the passes are not MFMA-bound: 2-2.5 TB/s of HBM at 4-5 waves per SIMD). The
// strip sine matrix Sst (KP x 16 NW) holds, in column 16 s + i, output i of strip s (Sst[j][16 s + i] = S[j][k] with
// k = 2(16 s + i) for an even strip, k = 2(16 (s - ne) + i) + 1 for an odd one; zero outside).
constexpr int SW = 16;

template <typename T>
__global__ void strip_sine_kernel(T* __restrict__ Sst, int64_t n, int64_t KP, int64_t LDS, int64_t ne) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= KP * LDS) return;
  const int64_t j = i / LDS, col = i % LDS, s = col / SW;
  const int64_t hp = (n + 1) / 2, h = n / 2;
  const bool odd = s >= ne;
  const int64_t c = (odd ? s - ne : s) * SW + col % SW, k = odd ? 2 * c + 1 : 2 * c;
  if (odd ? (j >= h || c >= h) : (j >= hp || c >= hp)) {
    Sst[i] = (T)0;
    return;
  }
  const int64_t m = ((j + 1) * (k + 1)) % (2 * (n + 1));
  Sst[i] = (T)(sqrt(2.0 / (double)(n + 1)) * sinpi((double)m / (double)(n + 1)));
}

// x axis: Out[o][k] = sum_j X[o][j] S[j][k] for 64 rows o per workgroup. PER = slab elements per thread (a power of
// two with PER * blockDim.x >= 64 * FK).
template <typename T, typename TIN, typename TOUT, int PRE, int POST, int PER>
__global__ void __launch_bounds__(1024) dst_rows_strip_kernel(const TIN* __restrict__ X, const double* __restrict__ pre,
                                                              const T* __restrict__ Sst, int64_t LDS, int64_t ne,
                                                              int64_t nstrip, int64_t O, int64_t n,
                                                              const double* __restrict__ post, double sign,
                                                              TOUT* __restrict__ Y) {
  static_assert(POST != POST_DIV, "the spectral division is fused into a strided pass");
  typedef typename Acc<T>::type AT;
  const int NT = blockDim.x;
  __shared__ T As[2][2][FK][TM + 1];   // As[buffer][U/V][k][m]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool active = wave < nstrip;                   // wave-uniform: a padding wave only stages
  const int side = wave >= ne ? 1 : 0;                 // 0 even outputs (U), 1 odd outputs (V)
  const int64_t c0 = SW * (wave - side * ne);
  const int64_t m0 = xcd_block(blockIdx.x, gridDim.x) * TM;
  const int64_t h = n / 2, hp = (n + 1) / 2;
  AT acc[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) acc[a] = AT{0, 0, 0, 0};
  const int li = lane & 15, lk = lane >> 4;
  TIN xa[PER], xb[PER];
  double pa[PER], pb[PER];
  auto fetch = [&](int64_t k0) {
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int e = t * NT + threadIdx.x;
      const int ec = e < TM * FK ? e : 0;
      const int64_t row = clamp_index(m0 + ec / FK, O) * n;
      const int64_t j = clamp_index(k0 + ec % FK, hp);
      xa[t] = X[row + j];
      xb[t] = X[row + n - 1 - j];
      if (PRE == PRE_SCALE) {
        pa[t] = pre[row + j];
        pb[t] = pre[row + n - 1 - j];
      }
    }
  };
  fetch(0);
  int buf = 0;
  const int64_t scol = SW * (int64_t)(active ? wave : 0) + li;
  for (int64_t k0 = 0; k0 < hp; k0 += FK) {
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int e = t * NT + threadIdx.x;
      if (e < TM * FK) {
        const int r = e / FK, kk = e % FK;
        const int64_t j = k0 + kk;
        const T a = PRE == PRE_SCALE ? (T)(pa[t] * (double)xa[t]) : (T)xa[t];
        const T b = PRE == PRE_SCALE ? (T)(pb[t] * (double)xb[t]) : (T)xb[t];
        const bool live = m0 + r < O;
        As[buf][0][kk][r] = (live && j < h) ? a + b : ((live && j < hp) ? a : (T)0);
        As[buf][1][kk][r] = (live && j < h) ? a - b : (T)0;
      }
    }
    __syncthreads();
    T bf[FK / 4];
#pragma unroll
    for (int q = 0; q < FK / 4; ++q) bf[q] = Sst[(k0 + 4 * q + lk) * LDS + scol];
    fetch(k0 + FK);   // unconditional (clamped past the end) so the wait counts stay static
    __builtin_amdgcn_sched_barrier(0);
    if (active) {
#pragma unroll
      for (int q = 0; q < FK / 4; ++q) {
#pragma unroll
        for (int a = 0; a < 4; ++a) acc[a] = mfma(As[buf][side][4 * q + lk][16 * a + li], bf[q], acc[a]);
      }
    }
    buf ^= 1;
  }
  if (!active) return;
  const int64_t nout = side ? h : hp;
  const int64_t c = c0 + li;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    double pv[4];
    if (POST == POST_SCALE) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gm = m0 + 16 * a + acc_row<T>(lk, r);
        pv[r] = post ? post[clamp_index(gm, O) * n + clamp_index(2 * c + side, n)] : 1.0;
      }
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t gm = m0 + 16 * a + acc_row<T>(lk, r);
      if (gm < O && c < nout) {
        const int64_t idx = gm * n + 2 * c + side;
        if (POST == POST_SCALE) {
          const double v = (double)acc[a][r];
          Y[idx] = (TOUT)(sign * (post ? pv[r] * v : v));
        } else {
          Y[idx] = (TOUT)acc[a][r];
        }
      }
    }
  }
}

// strided axes: Out_o[k][i] = sum_j S[k][j] X_o[j][i] for 64 inner columns i per workgroup.
template <typename T, int POST, int PER>
__global__ void __launch_bounds__(1024) dst_cols_strip_kernel(const T* __restrict__ X, const T* __restrict__ Sst,
                                                              int64_t LDS, int64_t ne, int64_t nstrip, int64_t n,
                                                              int64_t I, const double* __restrict__ den,
                                                              T* __restrict__ Y) {
  typedef typename Acc<T>::type AT;
  const int NT = blockDim.x;
  __shared__ T Bs[2][2][FK][TN + 1];   // Bs[buffer][U/V][k][i]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool active = wave < nstrip;
  const int side = wave >= ne ? 1 : 0;
  const int64_t c0 = SW * (wave - side * ne);
  const int64_t nib = (I + TN - 1) / TN;
  const int64_t tile = xcd_block(blockIdx.x, gridDim.x);
  const int64_t o = tile / nib, n0 = (tile % nib) * TN;
  const T* Xo = X + o * n * I;
  T* Yo = Y + o * n * I;
  const int64_t h = n / 2, hp = (n + 1) / 2;
  AT acc[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) acc[b] = AT{0, 0, 0, 0};
  const int li = lane & 15, lk = lane >> 4;
  T xa[PER], xb[PER];
  auto fetch = [&](int64_t k0) {
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int e = t * NT + threadIdx.x;
      const int ec = e < FK * TN ? e : 0;
      const int64_t j = clamp_index(k0 + ec / TN, hp), col = clamp_index(n0 + ec % TN, I);
      xa[t] = Xo[j * I + col];
      xb[t] = Xo[(n - 1 - j) * I + col];
    }
  };
  fetch(0);
  int buf = 0;
  const int64_t scol = SW * (int64_t)(active ? wave : 0) + li;
  for (int64_t k0 = 0; k0 < hp; k0 += FK) {
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int e = t * NT + threadIdx.x;
      if (e < FK * TN) {
        const int kk = e / TN, cc = e % TN;
        const int64_t j = k0 + kk;
        const bool live = n0 + cc < I;
        Bs[buf][0][kk][cc] = (live && j < h) ? xa[t] + xb[t] : ((live && j < hp) ? xa[t] : (T)0);
        Bs[buf][1][kk][cc] = (live && j < h) ? xa[t] - xb[t] : (T)0;
      }
    }
    __syncthreads();
    T af[FK / 4];
#pragma unroll
    for (int q = 0; q < FK / 4; ++q) af[q] = Sst[(k0 + 4 * q + lk) * LDS + scol];
    fetch(k0 + FK);   // unconditional (clamped past the end) so the wait counts stay static
    __builtin_amdgcn_sched_barrier(0);
    if (active) {
#pragma unroll
      for (int q = 0; q < FK / 4; ++q) {
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[b] = mfma(af[q], Bs[buf][side][4 * q + lk][16 * b + li], acc[b]);
      }
    }
    buf ^= 1;
  }
  if (!active) return;
  const int64_t nout = side ? h : hp;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int64_t c = c0 + acc_row<T>(lk, r);
    T dv[4];
    if (POST == POST_DIV) {
#pragma unroll
      for (int b = 0; b < 4; ++b) dv[b] = (T)den[clamp_index(2 * c + side, n) * I + clamp_index(n0 + 16 * b + li, I)];
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int64_t gi = n0 + 16 * b + li;
      if (c < nout && gi < I) {
        T v = acc[b][r];
        if (POST == POST_DIV) v = v / dv[b];
        Yo[(2 * c + side) * I + gi] = v;
      }
    }
  }
}

// Layout probe: D = A(16x4) * B(4x16) with one MFMA of type T, A[i][k] = a[i*4+k],
// B[k][j] = b[k*16+j]; D[i][j] is written to d[i*16+j] through the lane mapping the kernels above
// assume (operands and acc_row). The tests compare d with A*B for both instruction types.
template <typename T>
__global__ void mfma_probe_kernel(const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ d) {
  typedef typename Acc<T>::type AT;
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  const T af = (T)a[li * 4 + lk];
  const T bf = (T)b[lk * 16 + li];
  AT acc = AT{0, 0, 0, 0};
  acc = mfma(af, bf, acc);
#pragma unroll
  for (int r = 0; r < 4; ++r) d[acc_row<T>(lk, r) * 16 + li] = (double)acc[r];
}

// Spectral denominator omega * (a + b * lambda) on the interior modes from its separable 1-D factors
// (poisson.py:16-30, 118-131): q1 metric lambda = (lx my) mz + (mx ly) mz + (mx my) lz, seven-point
// lambda = (lx + ly) + lz, products and sums in the reference's order.
__global__ void __launch_bounds__(256) poisson_denominator_kernel(const double* __restrict__ lx,
                                                                  const double* __restrict__ wx,
                                                                  const double* __restrict__ ly,
                                                                  const double* __restrict__ wy,
                                                                  const double* __restrict__ lz,
                                                                  const double* __restrict__ wz, int64_t mx,
                                                                  int64_t my, int64_t mz, int q1, double omega,
                                                                  double a, double b, double* __restrict__ den) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= mx * my * mz) return;
  const int64_t x = i % mx, y = (i / mx) % my, z = i / (mx * my);
  double lam;
  if (q1)
    lam = (lx[x] * wy[y] * wz[z] + wx[x] * ly[y] * wz[z]) + wx[x] * wy[y] * lz[z];
  else
    lam = (lx[x] + ly[y]) + lz[z];
  den[i] = omega * (a + b * lam);
}

}  // namespace seqvp

using namespace seqvp;

static inline int64_t cdiv64(int64_t a, int64_t b) { return (a + b - 1) / b; }
// Strip sine matrix of length n <= 256: ne = ceil(ceil(n/2)/16) even strips, ceil(floor(n/2)/16) odd strips,
// max(their sum, 2) waves per workgroup, KP = ceil(n/2) rounded up to FK rows, 16 columns per wave.
static inline int64_t fold_kp(int64_t n) { return FK * cdiv64(cdiv64(n, 2), FK); }
static inline int64_t strip_ne(int64_t n) { return cdiv64(cdiv64(n, 2), SW); }
static inline int64_t strip_count(int64_t n) { return strip_ne(n) + cdiv64(n / 2, SW); }
static inline int64_t strip_waves(int64_t n) { return strip_count(n) > 2 ? strip_count(n) : 2; }
static inline int64_t strip_ld(int64_t n) { return SW * strip_waves(n); }
// slab elements (64 x FK) per thread of a workgroup of strip_waves(n) waves, rounded up to a power of two
static inline int strip_per(int64_t n) {
  const int64_t need = cdiv64(TM * FK, 64 * strip_waves(n));
  return need <= 1 ? 1 : need <= 2 ? 2 : need <= 4 ? 4 : 8;
}
// Words of the stored sine matrix of length n (strip layout for n <= 256, plain otherwise).
static inline int64_t sine_words(int64_t n) { return n <= 256 ? fold_kp(n) * strip_ld(n) : n * n; }

size_t seqvp_dst_workspace_bytes(int64_t mx, int64_t my, int64_t mz, int64_t ncomp) {
  return (size_t)(ncomp * mx * my * mz + sine_words(mx) + sine_words(my) + sine_words(mz)) * sizeof(double);
}

extern "C" int seqvp_mfma_probe(const double* a, const double* b, int fp32, double* d, void* stream) {
  if (!a || !b || !d) return SEQVP_EINVAL;
  if (fp32)
    hipLaunchKernelGGL((mfma_probe_kernel<float>), dim3(1), dim3(64), 0, (hipStream_t)stream, a, b, d);
  else
    hipLaunchKernelGGL((mfma_probe_kernel<double>), dim3(1), dim3(64), 0, (hipStream_t)stream, a, b, d);
  return launch_status();
}


// One pass along x (rows of length n): the folded kernel for n <= 256 (S = the folded matrix, LDF columns), the
// 64x64 tiling with the plain n x n matrix otherwise.
template <typename T, typename TIN, typename TOUT, int PRE, int POST>
static void rows_pass(const TIN* X, const double* pre, const T* S, int64_t O, int64_t n, const double* post,
                      double sign, TOUT* Y, hipStream_t s) {
  if (n > 256) {
    hipLaunchKernelGGL((dst_rows_kernel<T, TIN, TOUT, PRE, POST>), dim3(cdiv64(O, TM) * cdiv64(n, TN)), dim3(256), 0, s,
                       X, pre, S, O, n, post, sign, Y);
    return;
  }
  const dim3 g(cdiv64(O, TM)), b(64 * strip_waves(n));
  const int64_t ld = strip_ld(n), ne = strip_ne(n), ns = strip_count(n);
  switch (strip_per(n)) {
    case 1:
      hipLaunchKernelGGL((dst_rows_strip_kernel<T, TIN, TOUT, PRE, POST, 1>), g, b, 0, s, X, pre, S, ld, ne, ns, O, n,
                         post, sign, Y);
      break;
    case 2:
      hipLaunchKernelGGL((dst_rows_strip_kernel<T, TIN, TOUT, PRE, POST, 2>), g, b, 0, s, X, pre, S, ld, ne, ns, O, n,
                         post, sign, Y);
      break;
    case 4:
      hipLaunchKernelGGL((dst_rows_strip_kernel<T, TIN, TOUT, PRE, POST, 4>), g, b, 0, s, X, pre, S, ld, ne, ns, O, n,
                         post, sign, Y);
      break;
    default:
      hipLaunchKernelGGL((dst_rows_strip_kernel<T, TIN, TOUT, PRE, POST, 8>), g, b, 0, s, X, pre, S, ld, ne, ns, O, n,
                         post, sign, Y);
  }
}

// One strided pass of length n over `batch` blocks of inner size I.
template <typename T, int POST>
static void cols_pass(const T* X, const T* S, int64_t n, int64_t I, int64_t batch, const double* den, T* Y,
                      hipStream_t s) {
  if (n > 256) {
    hipLaunchKernelGGL((dst_cols_kernel<T, POST>), dim3(cdiv64(I, TN) * cdiv64(n, TM) * batch), dim3(256), 0, s, X, S,
                       n, I, den, Y);
    return;
  }
  const dim3 g(cdiv64(I, TN) * batch), b(64 * strip_waves(n));
  const int64_t ld = strip_ld(n), ne = strip_ne(n), ns = strip_count(n);
  switch (strip_per(n)) {
    case 1:
      hipLaunchKernelGGL((dst_cols_strip_kernel<T, POST, 1>), g, b, 0, s, X, S, ld, ne, ns, n, I, den, Y);
      break;
    case 2:
      hipLaunchKernelGGL((dst_cols_strip_kernel<T, POST, 2>), g, b, 0, s, X, S, ld, ne, ns, n, I, den, Y);
      break;
    case 4:
      hipLaunchKernelGGL((dst_cols_strip_kernel<T, POST, 4>), g, b, 0, s, X, S, ld, ne, ns, n, I, den, Y);
      break;
    default:
      hipLaunchKernelGGL((dst_cols_strip_kernel<T, POST, 8>), g, b, 0, s, X, S, ld, ne, ns, n, I, den, Y);
  }
}

// The sine matrix of length n into S: the strip layout (fold_kp(n) x strip_ld(n)) for n <= 256, plain n x n otherwise.
template <typename T>
static void sine_matrix(T* S, int64_t n, hipStream_t s) {
  if (n > 256) {
    hipLaunchKernelGGL((sine_matrix_kernel<T>), dim3(cdiv64(n * n, 256)), dim3(256), 0, s, S, n, n);
  } else {
    const int64_t kp = fold_kp(n), ld = strip_ld(n);
    hipLaunchKernelGGL((strip_sine_kernel<T>), dim3(cdiv64(kp * ld, 256)), dim3(256), 0, s, S, n, kp, ld, strip_ne(n));
  }
}

// T = compute/intermediate type. Intermediates ping-pong between the workspace and `out` (which has
// room for ncomp*mx*my*mz doubles, so for T = float it holds them too).
template <typename T>
static int dst_poisson(const double* rhs, const double* den, const double* scale, int negate, int64_t mx, int64_t my,
                       int64_t mz, int64_t ncomp, double* out, void* ws, size_t ws_bytes, void* stream) {
  if (!rhs || !den || !out || mx < 1 || my < 1 || mz < 1 || ncomp < 1 || ncomp > 65535) return SEQVP_EINVAL;
  if (scale == out) return SEQVP_EINVAL;
  if (!ws || ws_bytes < seqvp_dst_workspace_bytes(mx, my, mz, ncomp)) return SEQVP_EWORKSPACE;
  if (cdiv64(mx, TN) * cdiv64(my, TM) * ncomp * mz > 0x7fffffff) return SEQVP_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  const int64_t total = ncomp * mx * my * mz;
  T* tmp = (T*)ws;
  T* Sx = tmp + total;
  T* Sy = Sx + sine_words(mx);
  T* Sz = Sy + sine_words(my);
  T* alt = (T*)out;
  sine_matrix<T>(Sx, mx, s);
  sine_matrix<T>(Sy, my, s);
  sine_matrix<T>(Sz, mz, s);
  const int64_t O = ncomp * mz * my;
  const double sign = negate ? -1.0 : 1.0;
  // forward x (pre-scale): rhs -> tmp
  if (scale)
    rows_pass<T, double, T, PRE_SCALE, POST_NONE>(rhs, scale, Sx, O, mx, nullptr, 1.0, tmp, s);
  else
    rows_pass<T, double, T, PRE_NONE, POST_NONE>(rhs, nullptr, Sx, O, mx, nullptr, 1.0, tmp, s);
  // forward y: tmp -> alt (batch ncomp*mz, inner mx)
  cols_pass<T, POST_NONE>(tmp, Sy, my, mx, ncomp * mz, nullptr, alt, s);
  // forward z + division by the spectrum: alt -> tmp (batch ncomp, inner my*mx)
  cols_pass<T, POST_DIV>(alt, Sz, mz, my * mx, ncomp, den, tmp, s);
  // inverse z: tmp -> alt
  cols_pass<T, POST_NONE>(tmp, Sz, mz, my * mx, ncomp, nullptr, alt, s);
  // inverse y: alt -> tmp
  cols_pass<T, POST_NONE>(alt, Sy, my, mx, ncomp * mz, nullptr, tmp, s);
  // inverse x + sign*(s*x) in binary64: tmp -> out
  rows_pass<T, T, double, PRE_NONE, POST_SCALE>(tmp, nullptr, Sx, O, mx, scale, sign, out, s);
  return launch_status();
}

extern "C" int seqvp_dst_poisson(const double* rhs, const double* den, const double* scale, int negate, int64_t mx,
                                 int64_t my, int64_t mz, int64_t ncomp, double* out, void* ws, size_t ws_bytes,
                                 void* stream) {
  return dst_poisson<double>(rhs, den, scale, negate, mx, my, mz, ncomp, out, ws, ws_bytes, stream);
}

extern "C" int seqvp_dst_poisson_f32(const double* rhs, const double* den, const double* scale, int negate, int64_t mx,
                                     int64_t my, int64_t mz, int64_t ncomp, double* out, void* ws, size_t ws_bytes,
                                     void* stream) {
  return dst_poisson<float>(rhs, den, scale, negate, mx, my, mz, ncomp, out, ws, ws_bytes, stream);
}

extern "C" int seqvp_poisson_denominator(const double* lx, const double* wx, const double* ly, const double* wy,
                                         const double* lz, const double* wz, int64_t mx, int64_t my, int64_t mz,
                                         int q1, double omega, double a, double b, double* den, void* stream) {
  if (!lx || !ly || !lz || !den || mx < 1 || my < 1 || mz < 1) return SEQVP_EINVAL;
  if (q1 && (!wx || !wy || !wz)) return SEQVP_EINVAL;
  const int64_t n = mx * my * mz;
  hipLaunchKernelGGL(poisson_denominator_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, lx, wx, ly,
                     wy, lz, wz, mx, my, mz, q1, omega, a, b, den);
  return launch_status();
}

// Separable linear map of a scalar node field (nz, ny, nx): out = (Mz (x) My (x) Mx) in, i.e. each axis line
// multiplied by that axis's dense matrix (row-major, M[i][j] = coefficient i of unit input j). mx_t is Mx
// TRANSPOSED (the x pass right-multiplies rows). Used for the not-a-knot cubic spline coefficients of an image
// (the inverse collocation matrices of make_interp_spline, interpolation.py:29-40). FP64 MFMA passes:
// x in -> out, y out -> ws, z ws -> out. ws holds nx*ny*nz doubles.
extern "C" int seqvp_axis_transform(const double* in, double* out, const double* mx_t, const double* my,
                                    const double* mz, int64_t nx, int64_t ny, int64_t nz, void* ws, size_t ws_bytes,
                                    void* stream) {
  if (!in || !out || !mx_t || !my || !mz || in == out || nx < 1 || ny < 1 || nz < 1) return SEQVP_EINVAL;
  if (!ws || ws_bytes < (size_t)(nx * ny * nz) * sizeof(double)) return SEQVP_EWORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  double* tmp = (double*)ws;
  const int64_t O = nz * ny;
  hipLaunchKernelGGL((dst_rows_kernel<double, double, double, PRE_NONE, POST_NONE>), dim3(cdiv64(O, TM) * cdiv64(nx, TN)),
                     dim3(256), 0, s, in, nullptr, mx_t, O, nx, nullptr, 1.0, out);
  hipLaunchKernelGGL((dst_cols_kernel<double, POST_NONE>), dim3(cdiv64(nx, TN) * cdiv64(ny, TM) * nz), dim3(256), 0, s,
                     out, my, ny, nx, nullptr, tmp);
  hipLaunchKernelGGL((dst_cols_kernel<double, POST_NONE>), dim3(cdiv64(ny * nx, TN) * cdiv64(nz, TM)), dim3(256), 0,
                     s, tmp, mz, nz, ny * nx, nullptr, out);
  return launch_status();
}
