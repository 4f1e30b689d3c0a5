// Q1 cell sweeps for gfx950: soft JD/curl and closed27 barrier energies, the energy gradient, the
// GGN diagonal, the T62 node JD measure and the T66 global first JD root.
//
// One cell per lane, 256-lane workgroups (4 wave64), block ids remapped per XCD. FP64 VALU bound:
// 8-27 kflop per cell against 72-216 B of node reads. Energies reduce in double-double (wave
// butterfly, then a fixed-order block and grid pass). Node gathers are deterministic: each cell
// writes its vertex contributions to a structure-of-arrays buffer loc[a*3+r][cell] (coalesced
// stores; x-neighbour pairs in one wave are added by a lane shuffle and stored once), and one lane
// per node adds them in a fixed order after the node base -- no atomics, so every result is
// bit-reproducible.
#include <algorithm>
#include <mutex>

#include "q1_common.h"
#include "jd_extension.h"
#include "abi_util.h"

namespace seqvp {

constexpr int BLOCK = 256;
// Threads of the single-workgroup reductions of block partials: more loads in flight at once (a partial array of a
// native sweep is 26,454 blocks; 1024 threads fold 26 each).
constexpr int RB = 1024;

// blocks of a one-lane-per-cell sweep over the grid's cells (the partials' component stride)
__device__ __forceinline__ int64_t cell_blocks(const Dims& d) { return (d.ncell + BLOCK - 1) / BLOCK; }

// ------------------------------------------------------------------------------------------------
// Block and grid reductions of NS double-double sums and NM NaN-propagating minima. A block's 2 NS + NM
// results go to out[0], out[cs], out[2 cs], ...: the sweeps store their partials component by component
// (cs = the number of blocks), so that the fold below reads each component with coalesced loads.
template <int NS, int NM, int NT = BLOCK>
__device__ void block_reduce_store(DD* s, double* m, double* __restrict__ out, int64_t cs = 1) {
  __shared__ double sh[NT / 64][2 * NS + NM + 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NS; ++q) s[q] = dd_wave_sum(s[q]);
#pragma unroll
  for (int q = 0; q < NM; ++q) m[q] = wave_nan_min(m[q]);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NS; ++q) { sh[wave][2 * q] = s[q].hi; sh[wave][2 * q + 1] = s[q].lo; }
#pragma unroll
    for (int q = 0; q < NM; ++q) sh[wave][2 * NS + q] = m[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      DD t{sh[0][2 * q], sh[0][2 * q + 1]};
      for (int w = 1; w < NT / 64; ++w) dd_add(t, DD{sh[w][2 * q], sh[w][2 * q + 1]});
      out[2 * q * cs] = t.hi;
      out[(2 * q + 1) * cs] = t.lo;
    }
#pragma unroll
    for (int q = 0; q < NM; ++q) {
      double t = sh[0][2 * NS + q];
      for (int w = 1; w < NT / 64; ++w) t = nan_min(t, sh[w][2 * NS + q]);
      out[(2 * NS + q) * cs] = t;
    }
  }
}

// One workgroup folds all block partials (component q of block b at partials[q * nblocks + b]) in a fixed
// order; out gets NS sums
// (hi + lo) then NM minima, or, with INTERLEAVE (NS == NM), sum0, min0, sum1, min1, ...
// Fold of folder t (0 <= t < RB): blocks t, t + RB, t + 2 RB, ... in that order; BATCH of them are loaded at a
// time (unconditionally, from clamped addresses) so their latencies overlap instead of serialising the loop. The
// batch size changes only how many loads are in flight, not the order of the additions.
template <int NS, int NM, int BATCH = 8>
__device__ __forceinline__ void fold_partials(const double* __restrict__ partials, int64_t nblocks, int64_t t, DD* s,
                                              double* m) {
  constexpr int STRIDE = 2 * NS + NM;
#pragma unroll
  for (int q = 0; q < NS; ++q) s[q] = DD{0.0, 0.0};
#pragma unroll
  for (int q = 0; q < NM; ++q) m[q] = __builtin_inf();
  for (int64_t b0 = t; b0 < nblocks; b0 += BATCH * RB) {
    double v[BATCH][STRIDE];
#pragma unroll
    for (int i = 0; i < BATCH; ++i) {
      const int64_t b = b0 + i * RB < nblocks ? b0 + i * RB : nblocks - 1;
#pragma unroll
      for (int q = 0; q < STRIDE; ++q) v[i][q] = partials[q * nblocks + b];
    }
    // with a deep batch (one wave per SIMD, a large register budget) every load of the batch is issued before the
    // first addition: the scheduler would otherwise sink each load to its use and wait on them a few at a time. At
    // RB lanes per workgroup the budget is 128 VGPRs, and the barrier would spill.
    if constexpr (BATCH > 8) __builtin_amdgcn_sched_barrier(0);
    // a block past the end leaves the accumulators as they are (selected, not branched)
#pragma unroll
    for (int i = 0; i < BATCH; ++i) {
      const bool in = b0 + i * RB < nblocks;
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        DD t = s[q];
        dd_add(t, DD{v[i][2 * q], v[i][2 * q + 1]});
        s[q].hi = in ? t.hi : s[q].hi;
        s[q].lo = in ? t.lo : s[q].lo;
      }
#pragma unroll
      for (int q = 0; q < NM; ++q) m[q] = in ? nan_min(m[q], v[i][2 * NS + q]) : m[q];
    }
  }
}

template <int NS, int NM, bool INTERLEAVE = false>
__device__ __forceinline__ void reduce_partials_body(const double* __restrict__ partials, int64_t nblocks,
                                                     double* __restrict__ out) {
  constexpr int STRIDE = 2 * NS + NM;
  DD s[NS];
  double m[NM > 0 ? NM : 1];
  fold_partials<NS, NM>(partials, nblocks, threadIdx.x, s, m);
  __shared__ double fin[STRIDE];
  block_reduce_store<NS, NM, RB>(s, m, fin);
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NS; ++q) out[INTERLEAVE ? 2 * q : q] = fin[2 * q] + fin[2 * q + 1];
#pragma unroll
    for (int q = 0; q < NM; ++q) out[INTERLEAVE ? 2 * q + 1 : NS + q] = fin[2 * NS + q];
  }
}

template <int NS, int NM, bool INTERLEAVE = false>
__global__ void __launch_bounds__(RB) reduce_partials_kernel(const double* __restrict__ partials, int64_t nblocks,
                                                                double* __restrict__ out) {
  reduce_partials_body<NS, NM, INTERLEAVE>(partials, nblocks, out);
}

// The trial sweep's two reductions (soft terms; the barrier's sums and minima), in two short launches that give the
// bits of reduce_partials_kernel: the RB folders of its workgroup run as 4 workgroups of RB / 4 lanes per array
// (grid (4, 2)), each wave adds its lanes as block_reduce_store does and stores its sum (stage: 2 arrays x RB / 64
// waves x 6 doubles), and one thread per array then adds the RB / 64 wave sums in wave order. The folds then run on
// eight CUs instead of one.
constexpr int TRIAL_STAGE_WAVES = RB / 64;
constexpr int64_t TRIAL_STAGE_DOUBLES = 2 * TRIAL_STAGE_WAVES * 6;
// byte offset of the stage in the trial's workspace: after the two partial arrays (nb x 6 doubles each) and the
// 256-byte slot of the running least J
inline size_t trial_stage_offset(int64_t nb) { return (size_t)nb * 12 * sizeof(double) + 256; }

template <int NS, int NM>
__device__ __forceinline__ void trial_stage_fold(const double* __restrict__ partials, int64_t nblocks,
                                                 double* __restrict__ stage) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  DD s[NS];
  double m[NM > 0 ? NM : 1];
  // one wave per SIMD here: the register file holds 32 blocks' partials at once, so at the native size (26 blocks
  // per folder) every load of a folder is in flight together
  fold_partials<NS, NM, 32>(partials, nblocks, t, s, m);
#pragma unroll
  for (int q = 0; q < NS; ++q) s[q] = dd_wave_sum(s[q]);
#pragma unroll
  for (int q = 0; q < NM; ++q) m[q] = wave_nan_min(m[q]);
  if ((threadIdx.x & 63) == 0) {
    double* w = stage + (t >> 6) * 6;
#pragma unroll
    for (int q = 0; q < NS; ++q) { w[2 * q] = s[q].hi; w[2 * q + 1] = s[q].lo; }
#pragma unroll
    for (int q = 0; q < NM; ++q) w[2 * NS + q] = m[q];
  }
}

__global__ void __launch_bounds__(RB / 4) __attribute__((amdgpu_waves_per_eu(1, 1))) reduce_trial_stage_kernel(const double* __restrict__ soft,
                                                                    const double* __restrict__ bar, int64_t nblocks,
                                                                    double* __restrict__ stage) {
  if (blockIdx.y == 0)
    trial_stage_fold<3, 0>(soft, nblocks, stage);
  else
    trial_stage_fold<2, 2>(bar, nblocks, stage + TRIAL_STAGE_WAVES * 6);
}

__global__ void reduce_trial_final_kernel(const double* __restrict__ stage, double* __restrict__ out) {
  // the wave sums come to LDS first (one load per lane), so the serial sums below wait on no global load
  __shared__ double st[TRIAL_STAGE_DOUBLES];
  for (int i = threadIdx.x; i < TRIAL_STAGE_DOUBLES; i += blockDim.x) st[i] = stage[i];
  __syncthreads();
  if (threadIdx.x == 0) {           // soft: three sums
    const double* w = st;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      DD t{w[2 * q], w[2 * q + 1]};
      for (int v = 1; v < TRIAL_STAGE_WAVES; ++v) dd_add(t, DD{w[v * 6 + 2 * q], w[v * 6 + 2 * q + 1]});
      out[q] = t.hi + t.lo;
    }
  } else if (threadIdx.x == 1) {    // barrier: sum, min, sum, min
    const double* w = st + TRIAL_STAGE_WAVES * 6;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      DD t{w[2 * q], w[2 * q + 1]};
      for (int v = 1; v < TRIAL_STAGE_WAVES; ++v) dd_add(t, DD{w[v * 6 + 2 * q], w[v * 6 + 2 * q + 1]});
      out[3 + 2 * q] = t.hi + t.lo;
      double mm = w[4 + q];
      for (int v = 1; v < TRIAL_STAGE_WAVES; ++v) mm = nan_min(mm, w[v * 6 + 4 + q]);
      out[3 + 2 * q + 1] = mm;
    }
  }
}

__device__ __forceinline__ void load_ih(const double* __restrict__ k, double ih[3]) {
  ih[0] = k[K_IH];
  ih[1] = k[K_IH + 1];
  ih[2] = k[K_IH + 2];
}

__device__ __forceinline__ void gauss_point(int p, double s[3]) {
  s[0] = GAUSS_S[p % 3];
  s[1] = GAUSS_S[(p / 3) % 3];
  s[2] = GAUSS_S[p / 9];
}

__device__ __forceinline__ void closed_point(int p, double s[3]) {
  s[0] = CLOSED_S[p % 3];
  s[1] = CLOSED_S[(p / 3) % 3];
  s[2] = CLOSED_S[p / 9];
}

// ------------------------------------------------------------------------------------------------
// Soft JD/curl change energy (residuals.py:166-172) and its stable difference against a base
// (energy_comparison.py:196-200, blocks as in residuals.py:96-103).
template <bool WITH_BASE>
__global__ void __launch_bounds__(BLOCK, 2) soft_energy_kernel(const double* __restrict__ u, const double* __restrict__ ref,
                                                            const double* __restrict__ base,
                                                            const double* __restrict__ k, Dims d,
                                                            double* __restrict__ partials) {
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t c = blk * BLOCK + threadIdx.x;
  DD s[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  double none[1] = {0.0};
  // the reference field's edges live in LDS (register pressure: two fields' edges stay in VGPRs)
  __shared__ double er_lds[36 * BLOCK];
  lds_double* er = (lds_double*)(er_lds + threadIdx.x);
  if (c < d.ncell) {
    double ih[3];
    load_ih(k, ih);
    const double mu = k[K_MU], nu = k[K_NU];
    const int64_t b = cell_base_node(c, d);
    Edges Eu, Eb;
    {
      double v[8][3];
      load_cell(u, b, d, v);
      make_edges(v, ih, Eu);
      load_cell(ref, b, d, v);
      stage_edges(v, ih, er, BLOCK);
      if (WITH_BASE) {
        load_cell(base, b, d, v);
        make_edges(v, ih, Eb);
      }
    }
    double cj = 0.0, cc = 0.0;
    // points p = 3 row + ix in order; the three points of an x row share each field's x column
#pragma unroll 1
    for (int row = 0; row < 9; ++row) {
      const double(&wr)[4] = GAUSS_W.w[3 * row][0];
      double Ux[3], Rx[3], Bx[3];
      column_w(Eu, 0, wr, Ux);
      column_col_w(er, BLOCK, 0, wr, Rx);
      if (WITH_BASE) column_w(Eb, 0, wr, Bx);
#pragma unroll 1
      for (int p = 3 * row; p < 3 * row + 3; ++p) {
        const double qw = k[K_QW + p];
        double V[3][3], W[3][3], C[3], Cr[3];
        point_grad_yz_w(Eu, GAUSS_W.w[p], Ux, V);
        const double J = jacobian(V);
        curl_of(V, C);
        point_grad_col_yz_w(er, BLOCK, GAUSS_W.w[p], Rx, W);
        const double Jr = jacobian(W);
        curl_of(W, Cr);
        const double dj = J - Jr;
        const double d0 = C[0] - Cr[0], d1 = C[1] - Cr[1], d2 = C[2] - Cr[2];
        cj += qw * dj * dj;
        cc += (qw * d0 * d0 + qw * d1 * d1) + qw * d2 * d2;
        if (WITH_BASE) {
          // 1/2 w (y - x)(y + x) with y, x the trial and base residuals against the reference: the same stable
          // difference as with sqrt(w) folded into the residuals; the point's four terms are summed first and
          // added to the double-double accumulator once.
          double Y[3][3], Ca[3];
          point_grad_yz_w(Eb, GAUSS_W.w[p], Bx, Y);
          const double Ja = jacobian(Y);
          curl_of(Y, Ca);
          const double wm = 0.5 * mu * qw, wn = 0.5 * nu * qw;
          double t = wm * ((J - Ja) * (dj + (Ja - Jr)));
#pragma unroll
          for (int i = 0; i < 3; ++i) t += wn * ((C[i] - Ca[i]) * ((C[i] - Cr[i]) + (Ca[i] - Cr[i])));
          dd_add(s[2], t);
        }
      }
    }
    dd_add(s[0], 0.5 * mu * cj);
    dd_add(s[1], 0.5 * nu * cc);
  }
  block_reduce_store<3, 0>(s, none, partials + blk, cell_blocks(d));
}

// Doubles as unsigned keys in the same order (for atomic min/max).
__device__ __forceinline__ unsigned long long ordered_key(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double from_ordered_key(unsigned long long key) {
  return __longlong_as_double((long long)((key >> 63) ? (key & 0x7fffffffffffffffull) : ~key));
}

// cofactor rows of a plain 3x3 (no identity): R_r = M_{r+1} x M_{r+2} (geometry.py:11-23).
__device__ __forceinline__ void cofactor_plain(const double M[3][3], double A[3][3]) {
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const int j = (r + 1) % 3, kk = (r + 2) % 3;
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
      const int l = (cc + 1) % 3, m = (cc + 2) % 3;
      A[r][cc] = M[j][l] * M[kk][m] - M[j][m] * M[kk][l];
    }
  }
}

// A lower bound of J = det(I + grad u) over a cell's closed points, from its edges alone. Every point gradient is a
// blend of the four parallel edges with weights >= 0 summing to 1, so for any centre c_rj, |grad u(p)_rj - c_rj| <=
// max_k |e_rjk - c_rj| =: s_rj; with c the edges' mean, F(p) = Fc + E(p), Fc = I + c, |E(p)|_F <= S = |s|_F.
// Expanding det(Fc + E) like a cubic in E, J(p) >= det Fc - |cof Fc| S - |Fc| S^2/sqrt3 - S^3/(3 sqrt3) (Frobenius
// norms; |cof M|^2 = e2(sigma^2) <= |M|^4/3 and sigma1 sigma2 sigma3 <= (|M|^2/3)^(3/2)). nf: |Fc| + S >= |F(p)|.
// A nonfinite edge makes the bound or nf NaN or infinite, so every comparison a caller makes with it fails.
__device__ __forceinline__ double cell_j_lower_bound(const Edges& E, double& nf) {
  double F[3][3], S2 = 0.0;
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const double* e = E.e[j][r];
      const double m = 0.25 * ((e[0] + e[1]) + (e[2] + e[3]));
      double sp = 0.0;
#pragma unroll
      for (int k = 0; k < 4; ++k) sp = fmax(sp, fabs(e[k] - m));
      F[r][j] = (r == j ? 1.0 : 0.0) + m;
      S2 = fma(sp, sp, S2);
    }
  double A[3][3];
  cofactor_plain(F, A);
  const double det = (F[0][0] * A[0][0] + F[0][1] * A[0][1]) + F[0][2] * A[0][2];
  double nf2 = 0.0, na2 = 0.0;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      nf2 = fma(F[r][j], F[r][j], nf2);
      na2 = fma(A[r][j], A[r][j], na2);
    }
  const double inv_sqrt3 = 0.57735026918962584, inv_3sqrt3 = 0.19245008972987526;
  const double S = sqrt(S2), nfc = sqrt(nf2);
  nf = nfc + S;
  return det - (sqrt(na2) * S + nfc * S2 * inv_sqrt3 + S2 * S * inv_3sqrt3);
}

// cell_j_lower_bound less a margin (1e-12 (1 + nf^3)) that covers the rounding of the bound and of the sweeps' own J:
// every closed-point J a sweep computes for the cell is at least this value (NaN for a nonfinite edge).
__device__ __forceinline__ double cell_j_floor(const Edges& E) {
  double nf;
  const double jlb = cell_j_lower_bound(E, nf);
  return jlb - 1e-12 * (1.0 + nf * nf * nf);
}

// Whether every closed point of the cell has J >= level: a cell that passes has no point below the barrier
// activation, so its barrier terms (value, change, gradient, curvature) are exactly zero.
__device__ __forceinline__ bool cell_j_above(const Edges& E, double level) { return cell_j_floor(E) > level; }

// Closed27 barrier terms of one cell from its edges (and its base's): the minima of J (and base J) over the 27 points
// into m, the barrier value into s[0] and, against a base, the unweighted stable change into s[1]. QUIET: the
// caller knows the base's least closed27 J over the whole grid, base_min, and that it is >= act, so no base point is
// active and the first pass needs no base J: the wave's activity is its own points', and the base minimum is that
// value (the same bits: the base's J are the same function of the same edges wherever they are computed).
template <bool WITH_BASE, bool QUIET = false>
__device__ __forceinline__ void barrier_cell(Edges& Eu, Edges& Eb, const double* __restrict__ k, DD s[2], double m[2],
                                             double base_min = 0.0, bool clear = false) {
  static_assert(!QUIET || WITH_BASE, "a quiet base needs a base");
  const double w = k[K_W], delta = k[K_DELTA], act = k[K_ACT], scale = k[K_SCALE];
  // clear: the caller has shown from the edges that no closed point of the cell (nor of its base) is active and
  // that none can hold a minimum the sweep reports; a wave of clear cells adds nothing.
  if (__all(clear)) {
    if constexpr (QUIET) {
      m[1] = base_min;
      if (!(base_min >= act)) dd_add(s[0], __builtin_nan(""));
    }
    return;
  }
  // Minima over the 27 points, each with its compile-time closed blend, and whether any point is active.
  bool active = false;
  if constexpr (QUIET) {
    static_for<27>([&](auto P) {
      if constexpr (P % 9 == 0) opaque_edges(Eu);
      double V[3][3];
      closed_point_grad<P>(Eu, V);
      const double J = jacobian(V);
      m[0] = nan_min(m[0], J);
      active |= !(J >= act);
    });
    m[1] = base_min;
    if (!(base_min >= act)) dd_add(s[0], __builtin_nan(""));   // a broken guarantee gives a NaN barrier value
  } else {
    static_for<27>([&](auto P) {
      if constexpr (P % 9 == 0) {
        opaque_edges(Eu);
        if (WITH_BASE) opaque_edges(Eb);
      }
      double V[3][3];
      closed_point_grad<P>(Eu, V);
      const double J = jacobian(V);
      m[0] = nan_min(m[0], J);
      if (WITH_BASE) {
        double Y[3][3];
        closed_point_grad<P>(Eb, Y);
        const double Ja = jacobian(Y);
        m[1] = nan_min(m[1], Ja);
        active |= !(J >= act && Ja >= act);
      } else {
        active |= J < act;
      }
    });
  }
  // The barrier terms are zero at every point with J >= act (and, against a base, Ja >= act), so only a wave
  // with an active point evaluates them, over the points in order with the tabulated blends.
  if (__any(active)) {
    double val = 0.0;
    // points p = 3 row + ix in order; the three points of an x row share each field's x column
#pragma unroll 1
    for (int row = 0; row < 9; ++row) {
      const double(&wr)[4] = CLOSED_W.w[3 * row][0];
      double Ux[3], Bx[3];
      column_w(Eu, 0, wr, Ux);
      if (WITH_BASE) column_w(Eb, 0, wr, Bx);
#pragma unroll 1
      for (int p = 3 * row; p < 3 * row + 3; ++p) {
        double V[3][3];
        point_grad_yz_w(Eu, CLOSED_W.w[p], Ux, V);
        const double J = jacobian(V);
        val += barrier_value(J, w, delta, act, scale);
        if (WITH_BASE) {
          double Y[3][3];
          point_grad_yz_w(Eb, CLOSED_W.w[p], Bx, Y);
          const double Ja = jacobian(Y);
          if (!(J >= act && Ja >= act)) dd_add(s[1], barrier_change(Ja, J, delta, act, scale));
        }
      }
    }
    dd_add(s[0], val);
  }
}

// Closed27 barrier value, min J, unweighted stable change vs base, min base J
// (residuals.py:173-186; energy_comparison.py:234-240; triton_q1_kernels barrier_energy).
template <bool WITH_BASE>
__global__ void __launch_bounds__(BLOCK) barrier_energy_kernel(const double* __restrict__ u,
                                                               const double* __restrict__ base,
                                                               const double* __restrict__ k, Dims d,
                                                               double* __restrict__ partials) {
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t c = blk * BLOCK + threadIdx.x;
  DD s[2] = {{0.0, 0.0}, {0.0, 0.0}};
  double m[2] = {__builtin_inf(), __builtin_inf()};
  if (c < d.ncell) {
    double ih[3];
    load_ih(k, ih);
    const int64_t b = cell_base_node(c, d);
    Edges Eu, Eb;
    {
      double v[8][3];
      load_cell(u, b, d, v);
      make_edges(v, ih, Eu);
      if (WITH_BASE) {
        load_cell(base, b, d, v);
        make_edges(v, ih, Eb);
      }
    }
    barrier_cell<WITH_BASE>(Eu, Eb, k, s, m);
  }
  block_reduce_store<2, 2>(s, m, partials + blk, cell_blocks(d));
}

// The trial's soft and barrier energies in one sweep (the work of soft_energy_kernel and barrier_energy_kernel, in
// the same per-cell order and the same blocks, so their partials are the same bits): the vertex loads and the edges
// of the trial and its base are shared, and the closed-point pass runs on edges already in registers.
// One 256-cell block `blk` of the sweep (the single-pair kernel and the batched persistent kernel share it); er_lds
// is the workgroup's 36 * BLOCK doubles of LDS for the reference field's edges. BASE_IS_REF: the base is the
// reference itself (the outer comparison right after a self-reference refresh), so the base's Gauss-point J and
// curl are the reference's, already computed: the same values the base's own blends give, bit for bit.
template <bool WITH_BASE, bool BASE_IS_REF = false, bool QUIET = false>
__device__ __forceinline__ void trial_energy_block(const int64_t blk, const double* __restrict__ u,
                                                   const double* __restrict__ ref, const double* __restrict__ base,
                                                   const double* __restrict__ k, const Dims& d,
                                                   double* __restrict__ soft_partials,
                                                   double* __restrict__ barrier_partials, double* er_lds,
                                                   double base_min = 0.0,
                                                   unsigned long long* __restrict__ umin = nullptr) {
  const int64_t c = blk * BLOCK + threadIdx.x;
  // umin: the least closed27 J of u this launch's waves have computed so far (an ordered key; +inf at the start):
  // a cell whose J floor lies above it holds no minimum of u (nor does a base cell whose floor lies above the
  // base's known least J, base_min)
  const double ubound = umin ? from_ordered_key(__hip_atomic_load(umin, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                             : -__builtin_inf();
  bool clear = false;
  DD s[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  DD sb[2] = {{0.0, 0.0}, {0.0, 0.0}};
  double none[1] = {0.0};
  double m[2] = {__builtin_inf(), __builtin_inf()};
  lds_double* er = (lds_double*)(er_lds + threadIdx.x);
  if (c < d.ncell) {
    double ih[3];
    load_ih(k, ih);
    const double mu = k[K_MU], nu = k[K_NU];
    const int64_t b = cell_base_node(c, d);
    Edges Eu, Eb;
    {
      double v[8][3];
      load_cell(u, b, d, v);
      make_edges(v, ih, Eu);
      load_cell(ref, b, d, v);
      stage_edges(v, ih, er, BLOCK);
      if (WITH_BASE) {
        load_cell(base, b, d, v);
        make_edges(v, ih, Eb);
      }
    }
    if (umin) {
      const double act = k[K_ACT];
      const double fu = cell_j_floor(Eu);
      clear = fu > act && fu > ubound;
      if (WITH_BASE && !QUIET) {
        const double fb = cell_j_floor(Eb);
        clear = clear && fb > act && fb > base_min;   // base_min NaN (unknown): never clear
      }
    }
    double cj = 0.0, cc = 0.0;
#pragma unroll 1
    for (int row = 0; row < 9; ++row) {
      const double(&wr)[4] = GAUSS_W.w[3 * row][0];
      double Ux[3], Rx[3], Bx[3];
      column_w(Eu, 0, wr, Ux);
      column_col_w(er, BLOCK, 0, wr, Rx);
      if (WITH_BASE && !BASE_IS_REF) column_w(Eb, 0, wr, Bx);
#pragma unroll 1
      for (int p = 3 * row; p < 3 * row + 3; ++p) {
        const double qw = k[K_QW + p];
        double V[3][3], W[3][3], C[3], Cr[3];
        point_grad_yz_w(Eu, GAUSS_W.w[p], Ux, V);
        const double J = jacobian(V);
        curl_of(V, C);
        point_grad_col_yz_w(er, BLOCK, GAUSS_W.w[p], Rx, W);
        const double Jr = jacobian(W);
        curl_of(W, Cr);
        const double dj = J - Jr;
        const double d0 = C[0] - Cr[0], d1 = C[1] - Cr[1], d2 = C[2] - Cr[2];
        cj += qw * dj * dj;
        cc += (qw * d0 * d0 + qw * d1 * d1) + qw * d2 * d2;
        if (WITH_BASE) {
          double Ja, Ca[3];
          if (BASE_IS_REF) {
            Ja = Jr;
            Ca[0] = Cr[0];
            Ca[1] = Cr[1];
            Ca[2] = Cr[2];
          } else {
            double Y[3][3];
            point_grad_yz_w(Eb, GAUSS_W.w[p], Bx, Y);
            Ja = jacobian(Y);
            curl_of(Y, Ca);
          }
          const double wm = 0.5 * mu * qw, wn = 0.5 * nu * qw;
          double t = wm * ((J - Ja) * (dj + (Ja - Jr)));
#pragma unroll
          for (int i = 0; i < 3; ++i) t += wn * ((C[i] - Ca[i]) * ((C[i] - Cr[i]) + (Ca[i] - Cr[i])));
          dd_add(s[2], t);
        }
      }
    }
    dd_add(s[0], 0.5 * mu * cj);
    dd_add(s[1], 0.5 * nu * cc);
    barrier_cell<WITH_BASE, QUIET>(Eu, Eb, k, sb, m, base_min, clear);
  }
  if (umin) {
    const double wm = wave_nan_min(m[0]);
    if ((threadIdx.x & 63) == 0 && wm < __builtin_inf()) atomicMin(umin, ordered_key(wm));   // NaN: no update
  }
  block_reduce_store<3, 0>(s, none, soft_partials + blk, cell_blocks(d));
  block_reduce_store<2, 2>(sb, m, barrier_partials + blk, cell_blocks(d));
}

template <bool WITH_BASE, bool BASE_IS_REF = false, bool QUIET = false>
__global__ void __launch_bounds__(BLOCK, 2) trial_energy_kernel(const double* __restrict__ u,
                                                             const double* __restrict__ ref,
                                                             const double* __restrict__ base,
                                                             const double* __restrict__ k, Dims d,
                                                             double* __restrict__ soft_partials,
                                                             double* __restrict__ barrier_partials,
                                                             const double* __restrict__ base_min,
                                                             unsigned long long* __restrict__ umin) {
  __shared__ double er_lds[36 * BLOCK];
  trial_energy_block<WITH_BASE, BASE_IS_REF, QUIET>(xcd_block(blockIdx.x, gridDim.x), u, ref, base, k, d,
                                                    soft_partials, barrier_partials, er_lds,
                                                    base_min ? *base_min : __builtin_nan(""), umin);
}

__global__ void umin_init_kernel(unsigned long long* __restrict__ umin) { *umin = ordered_key(__builtin_inf()); }

// ---- batched trial energies: many pairs in one persistent launch ------------------------------------------------
// The work list is every pair's 256-cell blocks, `tile` consecutive blocks per work item; pair q's items start at
// item_begin (a multiple of 8). Workgroup w of a grid that is a multiple of 8 takes items w, w + grid, ...: on MI355X
// workgroups go to the 8 XCDs round robin, so all of w's items sit on XCD w mod 8 and the pair-local item index
// has the same residue; xcd_block then hands each XCD a contiguous range of the pair's cells, as in the single-pair
// launch. Each block writes the same partial slot with the same bits as trial_energy_kernel<true>, and the batched
// reduction folds them in the same order, so every pair's 7 results are those of seqvp_q1_trial_energy.
constexpr int TRIAL_BATCH_MAX = 32;
struct TrialBatchPair {
  const double* u;
  const double* ref;
  const double* base;
  const double* k;
  double* soft_partials;
  double* barrier_partials;
  double* out;
  int32_t nx, ny, nz;
  int32_t item_begin, nitems, nblocks;
};
struct TrialBatchTable {                 // passed by value (kernel arguments, < 4 KiB)
  int32_t npairs, tile, nitems, pad;
  TrialBatchPair p[TRIAL_BATCH_MAX];
};
static_assert(sizeof(TrialBatchTable) <= 4096, "kernel argument table");

__device__ __forceinline__ int batch_pair_of(const TrialBatchTable& t, int32_t item) {
  int q = 0;
  for (int i = 1; i < t.npairs; ++i) q = item >= t.p[i].item_begin ? i : q;
  return q;
}

__global__ void __launch_bounds__(BLOCK, 2) trial_energy_batch_kernel(const TrialBatchTable t) {
  __shared__ double er_lds[36 * BLOCK];
  for (int32_t item = blockIdx.x; item < t.nitems; item += gridDim.x) {
    const TrialBatchPair& P = t.p[batch_pair_of(t, item)];
    const int32_t li = item - P.item_begin;
    if (li >= P.nitems) continue;                 // alignment padding between pairs
    const Dims d = make_dims(P.nx, P.ny, P.nz);
    const int64_t r = xcd_block(li, P.nitems);
    for (int j = 0; j < t.tile; ++j) {
      const int64_t blk = r * t.tile + j;
      if (blk >= P.nblocks) break;
      __syncthreads();                            // the previous block's reduction has read its LDS
      trial_energy_block<true>(blk, P.u, P.ref, P.base, P.k, d, P.soft_partials, P.barrier_partials, er_lds);
    }
  }
}

__global__ void __launch_bounds__(RB) trial_energy_batch_reduce_kernel(const TrialBatchTable t) {
  const TrialBatchPair& P = t.p[blockIdx.x];
  if (blockIdx.y == 0)
    reduce_partials_body<3, 0>(P.soft_partials, P.nblocks, P.out);
  else
    reduce_partials_body<2, 2, true>(P.barrier_partials, P.nblocks, P.out + 3);
}

// ------------------------------------------------------------------------------------------------
// FP32 screening energies (the precision form; triton_q1_f32_kernels.py's soft_energy_f32 /
// barrier_energy_f32 roles): node values rounded to binary32 on load, every per-point term in binary32,
// per-lane Kahan-compensated binary32 sums, block and grid passes in double-double. Same outputs as the
// FP64 kernels; the error band of the screen is FP32_EPS*(64 (jd+curl) + 4096 barrier).
template <bool WITH_BASE>
__global__ void __launch_bounds__(BLOCK) soft_energy_f32_kernel(const double* __restrict__ u,
                                                                const double* __restrict__ ref,
                                                                const double* __restrict__ base,
                                                                const double* __restrict__ k, Dims d,
                                                                double* __restrict__ partials) {
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t c = blk * BLOCK + threadIdx.x;
  DD s[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  double none[1] = {0.0};
  if (c < d.ncell) {
    const float ih[3] = {(float)k[K_IH], (float)k[K_IH + 1], (float)k[K_IH + 2]};
    const float mu = (float)k[K_MU], nu = (float)k[K_NU];
    const int64_t b = cell_base_node(c, d);
    EdgesT<float> Eu, Er, Eb;
    {
      float v[8][3];
      load_cell(u, b, d, v);
      make_edges(v, ih, Eu);
      load_cell(ref, b, d, v);
      make_edges(v, ih, Er);
      if (WITH_BASE) {
        load_cell(base, b, d, v);
        make_edges(v, ih, Eb);
      }
    }
    KahanF cj{0.f, 0.f}, cc{0.f, 0.f}, cd{0.f, 0.f};
    // points p = 3 row + ix in order; the three points of an x row share each field's x column
#pragma unroll 1
    for (int row = 0; row < 9; ++row) {
      const float(&wr)[4] = GAUSS_WF.w[3 * row][0];
      float Ux[3], Rx[3], Bx[3];
      column_w(Eu, 0, wr, Ux);
      column_w(Er, 0, wr, Rx);
      if (WITH_BASE) column_w(Eb, 0, wr, Bx);
#pragma unroll 1
      for (int p = 3 * row; p < 3 * row + 3; ++p) {
        const float qw = (float)k[K_QW + p];
        float V[3][3], W[3][3], C[3], Cr[3];
        point_grad_yz_w(Eu, GAUSS_WF.w[p], Ux, V);
        const float J = jacobian(V);
        curl_of(V, C);
        point_grad_yz_w(Er, GAUSS_WF.w[p], Rx, W);
        const float Jr = jacobian(W);
        curl_of(W, Cr);
        const float dj = J - Jr;
        const float d0 = C[0] - Cr[0], d1 = C[1] - Cr[1], d2 = C[2] - Cr[2];
        kahan_add(cj, qw * dj * dj);
        kahan_add(cc, (qw * d0 * d0 + qw * d1 * d1) + qw * d2 * d2);
        if (WITH_BASE) {
          float Y[3][3], Ca[3];
          point_grad_yz_w(Eb, GAUSS_WF.w[p], Bx, Y);
          const float Ja = jacobian(Y);
          curl_of(Y, Ca);
          const float sm = sqrtf(mu * qw), sn = sqrtf(nu * qw);
          const float xj = sm * (Ja - Jr), yj = sm * dj;
          kahan_add(cd, 0.5f * ((yj - xj) * (yj + xj)));
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            const float xc = sn * (Ca[i] - Cr[i]), yc = sn * (C[i] - Cr[i]);
            kahan_add(cd, 0.5f * ((yc - xc) * (yc + xc)));
          }
        }
      }
    }
    dd_add(s[0], 0.5 * (double)mu * ((double)cj.s - (double)cj.c));
    dd_add(s[1], 0.5 * (double)nu * ((double)cc.s - (double)cc.c));
    dd_add(s[2], (double)cd.s - (double)cd.c);
  }
  block_reduce_store<3, 0>(s, none, partials + blk, cell_blocks(d));
}

template <bool WITH_BASE>
__global__ void __launch_bounds__(BLOCK) barrier_energy_f32_kernel(const double* __restrict__ u,
                                                                   const double* __restrict__ base,
                                                                   const double* __restrict__ k, Dims d,
                                                                   double* __restrict__ partials) {
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t c = blk * BLOCK + threadIdx.x;
  DD s[2] = {{0.0, 0.0}, {0.0, 0.0}};
  double m[2] = {__builtin_inf(), __builtin_inf()};
  if (c < d.ncell) {
    const float ih[3] = {(float)k[K_IH], (float)k[K_IH + 1], (float)k[K_IH + 2]};
    const float w = (float)k[K_W], delta = (float)k[K_DELTA], act = (float)k[K_ACT], scale = (float)k[K_SCALE];
    const int64_t b = cell_base_node(c, d);
    EdgesT<float> Eu, Eb;
    {
      float v[8][3];
      load_cell(u, b, d, v);
      make_edges(v, ih, Eu);
      if (WITH_BASE) {
        load_cell(base, b, d, v);
        make_edges(v, ih, Eb);
      }
    }
    KahanF val{0.f, 0.f}, chg{0.f, 0.f};
    float m0 = __builtin_inff(), m1 = __builtin_inff();
    // minima over the 27 points with their compile-time closed blends, and whether any point is active
    bool active = false;
    static_for<27>([&](auto P) {
      if constexpr (P % 9 == 0) {
        opaque_edges(Eu);
        if (WITH_BASE) opaque_edges(Eb);
      }
      float V[3][3];
      closed_point_grad<P>(Eu, V);
      const float J = jacobian(V);
      m0 = (J != J || m0 != m0) ? __builtin_nanf("") : (J < m0 ? J : m0);
      if (WITH_BASE) {
        float Y[3][3];
        closed_point_grad<P>(Eb, Y);
        const float Ja = jacobian(Y);
        m1 = (Ja != Ja || m1 != m1) ? __builtin_nanf("") : (Ja < m1 ? Ja : m1);
        active |= !(J >= act && Ja >= act);
      } else {
        active |= J < act;
      }
    });
    // the barrier sums are (0, 0) unless a point is active: a wave with one runs the full sequence of points
    if (__any(active)) {
#pragma unroll 1
      for (int p = 0; p < 27; ++p) {
        float V[3][3];
        point_grad_w(Eu, CLOSED_WF.w[p], V);
        const float J = jacobian(V);
        kahan_add(val, barrier_value(J, w, delta, act, scale));
        if (WITH_BASE) {
          float Y[3][3];
          point_grad_w(Eb, CLOSED_WF.w[p], Y);
          const float Ja = jacobian(Y);
          if (!(J >= act && Ja >= act)) kahan_add(chg, barrier_change(Ja, J, delta, act, scale));
        }
      }
    }
    dd_add(s[0], (double)val.s - (double)val.c);
    dd_add(s[1], (double)chg.s - (double)chg.c);
    m[0] = (double)m0;
    m[1] = (double)m1;
  }
  block_reduce_store<2, 2>(s, m, partials + blk, cell_blocks(d));
}

// ------------------------------------------------------------------------------------------------
// Per-cell vertex contributions (SoA: loc[(a*3+r)*ncell + c]).
//
// Gauss soft VJP: flux X = A*jd + curl adjoint with jd = mu qw (J-Jr), curl = nu qw (c-cr)
// (linearization.py:48-66,159-166); closed27 barrier VJP: X = A psi'(J) (linearization.py:169-177).
// The transpose accumulates G[j][r][k] += w_k(s) X[r][j] over points, then
// loc[a][r] = sum_j (+-1/h_j) G[j][r][k(a)]  (the adjoint of the edge-difference form).
struct EdgeAdj {
  double g[3][3][4];
};

__device__ __forceinline__ void adj_zero(EdgeAdj& G) {
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q) G.g[j][r][q] = 0.0;
}

// adj_accumulate with the weights of one table row.
__device__ __forceinline__ void adj_accumulate_w(EdgeAdj& G, const double (&w)[3][4], const double X[3][3]) {
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      G.g[j][r][0] += w[j][0] * X[r][j];
      G.g[j][r][1] += w[j][1] * X[r][j];
      G.g[j][r][2] += w[j][2] * X[r][j];
      G.g[j][r][3] += w[j][3] * X[r][j];
    }
}

// adj_accumulate_w for the y and z columns of X; the x column is summed over the points of an x row, which share
// its weights, and added once per row by adj_accumulate_x.
__device__ __forceinline__ void adj_accumulate_yz_w(EdgeAdj& G, const double (&w)[3][4], const double X[3][3],
                                                    double Xx[3]) {
#pragma unroll
  for (int r = 0; r < 3; ++r) Xx[r] += X[r][0];
#pragma unroll
  for (int j = 1; j < 3; ++j)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      G.g[j][r][0] += w[j][0] * X[r][j];
      G.g[j][r][1] += w[j][1] * X[r][j];
      G.g[j][r][2] += w[j][2] * X[r][j];
      G.g[j][r][3] += w[j][3] * X[r][j];
    }
}

__device__ __forceinline__ void adj_accumulate_x(EdgeAdj& G, const double (&wr)[4], const double Xx[3]) {
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    G.g[0][r][0] += wr[0] * Xx[r];
    G.g[0][r][1] += wr[1] * Xx[r];
    G.g[0][r][2] += wr[2] * Xx[r];
    G.g[0][r][3] += wr[3] * Xx[r];
  }
}

__device__ __forceinline__ void adj_accumulate(EdgeAdj& G, const double s[3], const double X[3][3]) {
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double so1 = s[other1(j)], so2 = s[other2(j)];
    const double w00 = (1.0 - so1) * (1.0 - so2), w10 = so1 * (1.0 - so2);
    const double w01 = (1.0 - so1) * so2, w11 = so1 * so2;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      G.g[j][r][0] += w00 * X[r][j];
      G.g[j][r][1] += w10 * X[r][j];
      G.g[j][r][2] += w01 * X[r][j];
      G.g[j][r][3] += w11 * X[r][j];
    }
  }
}

// Vertex a's k index for direction j: the bits of a on the two other axes.
__device__ __forceinline__ constexpr int vertex_k(int a, int j) {
  return ((a >> other1(j)) & 1) | (((a >> other2(j)) & 1) << 1);
}

// ---- per-cell vertex contributions ("locals") and their gathers -------------------------------------------------
// Vertex a + 1 (x bit 1) of cell c is vertex a (x bit 0) of cell c + 1 when c + 1 is in the same x row. When c + 1
// is also in the same wave (lane = c mod 64: the sweeps map cell blk*BLOCK + threadIdx.x) the pair is added there
// by a lane shuffle and stored once: plane a (x bit 0) of cell c holds v_a(c) + v_{a+1}(c - 1) where c merges, and
// plane a + 1 holds v_{a+1}(c) only for the cells whose successor does not merge (a wave's last lane, a row's last
// cell); its other entries are never written or read. Halves the locals traffic of the sweeps and gathers.
// Callers must map cell c to lane c & 63 (a block of a multiple of 64 lanes starting at a multiple of 64) and keep
// every lane of the wave converged at locals_store_pair (no per-lane early return before it): the shuffle reads the
// neighbouring lane's value, and a lane that left would leave a plane unwritten that the gather then reads.
static_assert(BLOCK % 64 == 0, "the locals merge assumes lane = cell & 63");
__device__ __forceinline__ bool locals_merge(int64_t c, int64_t x) { return (c & 63) != 0 && x != 0; }
__device__ __forceinline__ bool locals_keep_odd(int64_t c, int64_t x, const Dims& d) {
  return ((c + 1) & 63) == 0 || x == d.cx - 1;
}

// Stores the pair (a, a + 1) of component r; every lane of the wave calls it (the shuffle), `valid` lanes store.
__device__ __forceinline__ void locals_store_pair(double even, double odd, double* __restrict__ loc, int a, int r,
                                                  int64_t c, int64_t x, bool valid, const Dims& d) {
  const double prev = __shfl_up(odd, 1, 64);
  const double m = locals_merge(c, x) ? even + prev : even;
  if (valid) {
    loc[(int64_t)(a * 3 + r) * d.ncell + c] = m;
    if (locals_keep_odd(c, x, d)) loc[(int64_t)((a + 1) * 3 + r) * d.ncell + c] = odd;
  }
}

// Adds the contributions of the cells around node (x, y, z) to acc in a fixed order: for each (y, z) offset of
// the four pairs, the +x cell's plane (merged with the -x cell where it merges), then the -x cell's own plane
// where it was not merged. EDGE: the node may lie on the boundary (cells out of range are skipped).
template <bool EDGE>
__device__ __forceinline__ double locals_gather(const double* __restrict__ loc, int r, int64_t x, int64_t y, int64_t z,
                                                const Dims& d, double acc) {
#pragma unroll
  for (int pr = 0; pr < 4; ++pr) {
    const int64_t ya = y - (pr & 1), za = z - (pr >> 1);
    if (EDGE && (ya < 0 || ya >= d.cy || za < 0 || za >= d.cz)) continue;
    const int a = 2 * pr;
    const int64_t ce = (za * d.cy + ya) * d.cx + x;   // the +x cell (x bit 0 vertex); its -x neighbour is ce - 1
    const bool has_e = !EDGE || x < d.cx, has_o = !EDGE || x > 0;
    bool merged = false;
    if (has_e) {
      acc += loc[(int64_t)(a * 3 + r) * d.ncell + ce];
      merged = locals_merge(ce, x);
    }
    if (has_o && !merged) acc += loc[(int64_t)((a + 1) * 3 + r) * d.ncell + ce - 1];
  }
  return acc;
}

__device__ __forceinline__ void adj_store(const EdgeAdj& G, const double ih[3], double* __restrict__ loc, int64_t c,
                                          int64_t x, bool valid, const Dims& d) {
#pragma unroll
  for (int a = 0; a < 8; a += 2)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      double v[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int b = a + q;
        double t0 = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const double t = ih[j] * G.g[j][r][vertex_k(b, j)];
          t0 = ((b >> j) & 1) ? t0 + t : t0 - t;
        }
        v[q] = t0;
      }
      locals_store_pair(v[0], v[1], loc, a, r, c, x, valid, d);
    }
}

__device__ __forceinline__ void soft_vjp(const Edges& Eu, const lds_double* er, int stride, const double* __restrict__ k,
                                         EdgeAdj& G) {
  const double mu = k[K_MU], nu = k[K_NU];
  // points p = 3 row + ix in order; the three points of an x row share the x columns (and their weights)
#pragma unroll 1
  for (int row = 0; row < 9; ++row) {
    const double(&wr)[4] = GAUSS_W.w[3 * row][0];
    double Ux[3], Rx[3], Xx[3] = {0.0, 0.0, 0.0};
    column_w(Eu, 0, wr, Ux);
    column_col_w(er, stride, 0, wr, Rx);
#pragma unroll 1
    for (int p = 3 * row; p < 3 * row + 3; ++p) {
      const double qw = k[K_QW + p];
      double V[3][3], A[3][3], W[3][3], C[3], Cr[3];
      point_grad_yz_w(Eu, GAUSS_W.w[p], Ux, V);
      const double J = cofactor_jacobian(V, A);
      curl_of(V, C);
      point_grad_col_yz_w(er, stride, GAUSS_W.w[p], Rx, W);
      const double Jr = jacobian(W);
      curl_of(W, Cr);
      const double jd = mu * qw * (J - Jr);
      const double c0 = nu * qw * (C[0] - Cr[0]), c1 = nu * qw * (C[1] - Cr[1]), c2 = nu * qw * (C[2] - Cr[2]);
      double X[3][3];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int j = 0; j < 3; ++j) X[r][j] = A[r][j] * jd;
      X[2][1] += c0;
      X[1][2] -= c0;
      X[0][2] += c1;
      X[2][0] -= c1;
      X[1][0] += c2;
      X[0][1] -= c2;
      adj_accumulate_yz_w(G, GAUSS_W.w[p], X, Xx);
    }
    adj_accumulate_x(G, wr, Xx);
  }
}

// quiet: the cell's edges bound J at or above the activation at every closed point (cell_j_above).
__device__ __forceinline__ void barrier_vjp(Edges& Eu, const double* __restrict__ k, EdgeAdj& G, bool quiet) {
  const double w = k[K_W], delta = k[K_DELTA], act = k[K_ACT], scale = k[K_SCALE];
  // psi'(J) is exactly 0 at and above the activation: a wave whose cells are all quiet contributes nothing; else J
  // at the 27 points with their compile-time closed blends first, and a wave with no point below it contributes
  // nothing either.
  if (__all(quiet)) return;
  bool active = false;
  static_for<27>([&](auto P) {
    if constexpr (P % 9 == 0) opaque_edges(Eu);
    double V[3][3];
    closed_point_grad<P>(Eu, V);
    active |= jacobian(V) < act;
  });
  if (!__any(active)) return;
  // points p = 3 row + ix in order; the three points of an x row share the x columns (and their weights)
#pragma unroll 1
  for (int row = 0; row < 9; ++row) {
    const double(&wr)[4] = CLOSED_W.w[3 * row][0];
    double Ux[3], Xx[3] = {0.0, 0.0, 0.0};
    column_w(Eu, 0, wr, Ux);
    bool any = false;
#pragma unroll 1
    for (int p = 3 * row; p < 3 * row + 3; ++p) {
      double V[3][3];
      point_grad_yz_w(Eu, CLOSED_W.w[p], Ux, V);
      // psi'(J) is exactly 0 at and above the activation: a point where every lane of the wave is
      // inactive contributes nothing and is skipped (wave-uniform branch).
      if (!__any(jacobian(V) < act)) continue;
      any = true;
      double A[3][3];
      const double J = cofactor_jacobian(V, A);
      const double b1 = barrier_first(J, w, delta, act, scale);
      double X[3][3];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int j = 0; j < 3; ++j) X[r][j] = A[r][j] * b1;
      adj_accumulate_yz_w(G, CLOSED_W.w[p], X, Xx);
    }
    if (any) adj_accumulate_x(G, wr, Xx);
  }
}

// MODE 0: soft (+ barrier) into one buffer; MODE 1: barrier only; MODE 2: soft into loc, barrier into
// loc2 (the joint GF/GE form, linearization.py:217-238).
template <int MODE, bool SOFT, bool BARRIER>
__global__ void __launch_bounds__(BLOCK, 2) grad_locals_kernel(const double* __restrict__ u, const double* __restrict__ ref,
                                                            const double* __restrict__ k, Dims d,
                                                            double* __restrict__ loc, double* __restrict__ loc2) {
  // every lane runs to the end (the locals' lane shuffle); lanes past the last cell work on the last cell and
  // store nothing
  const int64_t c = xcd_block(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
  const bool valid = c < d.ncell;
  const int64_t cv = valid ? c : d.ncell - 1;
  const int64_t xc = cv % d.cx;
  double ih[3];
  load_ih(k, ih);
  const int64_t b = cell_base_node(cv, d);
  Edges Eu;
  {
    double v[8][3];
    load_cell(u, b, d, v);
    make_edges(v, ih, Eu);
  }
  // whether the barrier terms vanish on the whole cell, from its edges (before the soft pass, while few registers
  // are live)
  const bool quiet = BARRIER && cell_j_above(Eu, k[K_ACT]);
  EdgeAdj G;
  adj_zero(G);
  if (SOFT) {
    // the reference field's edges live in LDS (register pressure)
    __shared__ double er_lds[36 * BLOCK];
    lds_double* er = (lds_double*)(er_lds + threadIdx.x);
    double v[8][3];
    load_cell(ref, b, d, v);
    stage_edges(v, ih, er, BLOCK);
    soft_vjp(Eu, er, BLOCK, k, G);
  }
  if (MODE == 2) {
    adj_store(G, ih, loc, c, xc, valid, d);
    adj_zero(G);
  }
  if (BARRIER) barrier_vjp(Eu, k, G, quiet);
  adj_store(G, ih, MODE == 2 ? loc2 : loc, c, xc, valid, d);
}

// GGN diagonal per cell (linearization.py:357-380; gen_triton_q1.py:447-515):
// sum_Gauss qw (mu (A_r . dN_a)^2 + nu sum_{j!=r} dN_a,j^2) + sum_closed27 psi''(J) (A_r . dN_a)^2.
// dN_a,j(s) = (+-1/h_j) f(a_o1, s_o1) f(a_o2, s_o2) (cell_q1.py:118-127).
__device__ __forceinline__ void basis_grad(int a, const double s[3], const double ih[3], double dn[3]) {
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int o1 = other1(j), o2 = other2(j);
    const double f1 = ((a >> o1) & 1) ? s[o1] : 1.0 - s[o1];
    const double f2 = ((a >> o2) & 1) ? s[o2] : 1.0 - s[o2];
    const double t = ih[j] * f1 * f2;
    dn[j] = ((a >> j) & 1) ? t : -t;
  }
}

__global__ void __launch_bounds__(BLOCK) ggn_locals_kernel(const double* __restrict__ u, const double* __restrict__ k,
                                                           Dims d, double* __restrict__ loc) {
  // every lane runs to the end (the locals' lane shuffle); lanes past the last cell store nothing
  const int64_t c = xcd_block(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
  const bool valid = c < d.ncell;
  const int64_t cv = valid ? c : d.ncell - 1;
  double ih[3];
  load_ih(k, ih);
  const double mu = k[K_MU], nu = k[K_NU];
  const double w = k[K_W], delta = k[K_DELTA], act = k[K_ACT], scale = k[K_SCALE];
  const int64_t b = cell_base_node(cv, d);
  Edges Eu;
  {
    double v[8][3];
    load_cell(u, b, d, v);
    make_edges(v, ih, Eu);
  }
  const bool quiet = cell_j_above(Eu, act);
  // The curl part nu sum_Gauss qw sum_{j != r} dN_a,j^2 does not depend on u: 24 workgroup constants, one lane each
  // (every lane of the workgroup reaches the barrier)
  __shared__ double curl_part[24];
  if (threadIdx.x < 24) {
    const int a = threadIdx.x / 3, r = threadIdx.x % 3;
    double t = 0.0;
    for (int p = 0; p < 27; ++p) {
      double sp[3], dn[3];
      gauss_point(p, sp);
      basis_grad(a, sp, ih, dn);
      const double sq0 = dn[0] * dn[0], sq1 = dn[1] * dn[1], sq2 = dn[2] * dn[2];
      const double cs = r == 0 ? sq1 + sq2 : (r == 1 ? sq0 + sq2 : sq0 + sq1);
      t += k[K_QW + p] * cs;
    }
    curl_part[threadIdx.x] = nu * t;
  }
  __syncthreads();
  double acc[8][3];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int r = 0; r < 3; ++r) acc[a][r] = 0.0;
  // the JD part mu sum_Gauss qw (A_r . dN_a)^2
#pragma unroll 1
  for (int p = 0; p < 27; ++p) {
    double sp[3];
    gauss_point(p, sp);
    const double qm = k[K_QW + p] * mu;
    double V[3][3], A[3][3];
    point_grad_w(Eu, GAUSS_W.w[p], V);
    cofactor_jacobian(V, A);
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      double dn[3];
      basis_grad(a, sp, ih, dn);
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const double l = (A[r][0] * dn[0] + A[r][1] * dn[1]) + A[r][2] * dn[2];
        acc[a][r] = fma(qm * l, l, acc[a][r]);
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int r = 0; r < 3; ++r) acc[a][r] += curl_part[3 * a + r];
  // psi''(J) is exactly 0 at and above the activation: a wave whose cells' edges bound J at or above it at every
  // closed point (cell_j_above) skips the barrier part; else J at the 27 points with their compile-time closed
  // blends first, and a wave with no point below it skips it too.
  bool active = false;
  if (!__all(quiet)) static_for<27>([&](auto P) {
    if constexpr (P % 9 == 0) opaque_edges(Eu);
    double V[3][3];
    closed_point_grad<P>(Eu, V);
    active |= jacobian(V) < act;
  });
#pragma unroll 1
  for (int p = 0; p < 27 && __any(active); ++p) {
    double sp[3];
    closed_point(p, sp);
    double V[3][3], A[3][3];
    point_grad_w(Eu, CLOSED_W.w[p], V);
    const double J = cofactor_jacobian(V, A);
    const double b2 = barrier_second(J, w, delta, act, scale);
    if (b2 != 0.0) {
#pragma unroll
      for (int a = 0; a < 8; ++a) {
        double dn[3];
        basis_grad(a, sp, ih, dn);
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          const double l = (A[r][0] * dn[0] + A[r][1] * dn[1]) + A[r][2] * dn[2];
          acc[a][r] += b2 * (l * l);
        }
      }
    }
  }
  const int64_t xc = cv % d.cx;
#pragma unroll
  for (int a = 0; a < 8; a += 2)
#pragma unroll
    for (int r = 0; r < 3; ++r) locals_store_pair(acc[a][r], acc[a + 1][r], loc, a, r, c, xc, valid, d);
}

// out[n,r] = base[n,r] + the cell contributions of loc (then of loc2) to node n in locals_gather's fixed order
// (the reference gathers vertex by vertex, gen_triton_q1.py:425-444). base may be NULL.
template <bool TWO>
__global__ void __launch_bounds__(BLOCK) gather_kernel(const double* __restrict__ base,
                                                       const double* __restrict__ loc,
                                                       const double* __restrict__ loc2, Dims d,
                                                       double* __restrict__ out) {
  const int64_t nnode = d.nx * d.ny * d.nz;
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= nnode) return;
  const int64_t x = i % d.nx, y = (i / d.nx) % d.ny, z = i / (d.nx * d.ny);
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    double acc = base ? base[3 * i + r] : 0.0;
    acc = locals_gather<true>(loc, r, x, y, z, d, acc);
    if (TWO) acc = locals_gather<true>(loc2, r, x, y, z, d, acc);
    out[3 * i + r] = acc;
  }
}

// ------------------------------------------------------------------------------------------------
// T62 cell minima: min(closed27 J, exact edge-quadratic minima) (jd_affine.py:17-48).
__device__ __forceinline__ double edge_quadratic_min(double f0, double f1, double f2) {
  const double a = 2.0 * (f0 - 2.0 * f1 + f2);
  const double b = -3.0 * f0 + 4.0 * f1 - f2;
  const bool inside = (a > 0.0) && (b < 0.0) && (-b < 2.0 * a);
  return inside ? f0 - b * b / (4.0 * a) : __builtin_inf();
}

// The 27 closed-point Jacobians of cell c (index 9 iz + 3 iy + ix).
__device__ __forceinline__ void cell_closed_jd(const double* __restrict__ u, const double ih[3], const Dims& d,
                                               int64_t c, double J[27]) {
  Edges Eu;
  {
    double v[8][3];
    load_cell(u, cell_base_node(c, d), d, v);
    make_edges(v, ih, Eu);
  }
  static_for<27>([&](auto P) {
    double V[3][3];
    closed_point_grad<P>(Eu, V);
    J[P] = jacobian(V);
  });
}

// m_c per cell. With EXT, a cell whose Bernstein hull reaches below min(m_c, activation) is appended to the
// candidate list of the T69 extension (jd_affine.py:154-157; counters[0] counts them, list order is free).
template <bool EXT>
__global__ void __launch_bounds__(BLOCK) cell_min_jd_kernel(const double* __restrict__ u, const double* __restrict__ k,
                                                            Dims d, double activation, double* __restrict__ minima,
                                                            uint32_t* __restrict__ list,
                                                            unsigned long long* __restrict__ counters) {
  const int64_t c = xcd_block(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
  if (c >= d.ncell) return;
  double ih[3];
  load_ih(k, ih);
  double J[27];
  cell_closed_jd(u, ih, d, c, J);
  double m = J[0];
#pragma unroll
  for (int p = 1; p < 27; ++p) m = fmin(m, J[p]);
  // 12 edges through the closed27 points (jd_affine.py:17-21): x-lines, y-lines, z-lines.
#pragma unroll
  for (int iz = 0; iz < 3; iz += 2)
#pragma unroll
    for (int iy = 0; iy < 3; iy += 2) m = fmin(m, edge_quadratic_min(J[9 * iz + 3 * iy], J[9 * iz + 3 * iy + 1], J[9 * iz + 3 * iy + 2]));
#pragma unroll
  for (int iz = 0; iz < 3; iz += 2)
#pragma unroll
    for (int ix = 0; ix < 3; ix += 2) m = fmin(m, edge_quadratic_min(J[9 * iz + ix], J[9 * iz + 3 + ix], J[9 * iz + 6 + ix]));
#pragma unroll
  for (int iy = 0; iy < 3; iy += 2)
#pragma unroll
    for (int ix = 0; ix < 3; ix += 2) m = fmin(m, edge_quadratic_min(J[3 * iy + ix], J[9 + 3 * iy + ix], J[18 + 3 * iy + ix]));
  minima[c] = m;
  if (EXT && t69::bernstein_min_inplace(J) < fmin(m, activation)) list[atomicAdd(&counters[0], 1ull)] = (uint32_t)c;
}

// Order-preserving map of doubles to unsigned keys (for an integer atomic min).
__global__ void jd_extension_init_kernel(unsigned long long* __restrict__ counters) {
  if (threadIdx.x < 3) counters[threadIdx.x] = 0ull;
  if (threadIdx.x == 3) counters[3] = ordered_key(__builtin_inf());
}

// m_c <- min(m_c, mu_c) on the candidate cells (jd_affine.py:176-188), one lane per candidate. counters[1..3]
// collect the record: cells lowered by more than 0.1 %, cells lowered at all, the least mu.
constexpr int EXT_BLOCK = 64;
__global__ void __launch_bounds__(EXT_BLOCK) jd_extension_kernel(const double* __restrict__ u,
                                                                 const double* __restrict__ k, Dims d,
                                                                 const uint32_t* __restrict__ list,
                                                                 unsigned long long* __restrict__ counters,
                                                                 double* __restrict__ minima) {
  const unsigned long long n = counters[0];
  double ih[3];
  load_ih(k, ih);
  for (unsigned long long i = (unsigned long long)blockIdx.x * EXT_BLOCK + threadIdx.x; i < n;
       i += (unsigned long long)gridDim.x * EXT_BLOCK) {
    const int64_t c = list[i];
    double J[27];
    cell_closed_jd(u, ih, d, c, J);
    const double mu = t69::face_interior_min(J);
    const double m = minima[c];
    if (mu < m * (1.0 - 1e-3)) atomicAdd(&counters[1], 1ull);
    if (mu < m) atomicAdd(&counters[2], 1ull);
    atomicMin(&counters[3], ordered_key(mu));
    minima[c] = nan_min(m, mu);
  }
}

__global__ void jd_extension_stats_kernel(const unsigned long long* __restrict__ counters, double* __restrict__ stats) {
  if (threadIdx.x != 0) return;
  stats[0] = (double)counters[0];
  stats[1] = (double)counters[1];
  stats[2] = (double)counters[2];
  stats[3] = from_ordered_key(counters[3]);
}

// ------------------------------------------------------------------------------------------------
// JD witnesses (witness_screen.py:22-88; jd_step.py:13-17, 177-239).
//
// Per tile of `tile` consecutive cells: the least closed27 J (a nonfinite J counts as +inf), the first flat
// index (cell - tile start) * 27 + p where it occurs, and the 24 vertex values of that cell, so the host can
// check the nominated point with an interval enclosure. out: 26 doubles per tile.
__global__ void __launch_bounds__(BLOCK) closed27_tile_minima_kernel(const double* __restrict__ u,
                                                                     const double* __restrict__ k, Dims d,
                                                                     int64_t tile, double* __restrict__ out) {
  const int64_t t = blockIdx.x;
  const int64_t start = t * tile, end = start + tile < d.ncell ? start + tile : d.ncell;
  double ih[3];
  load_ih(k, ih);
  double best = __builtin_inf();
  int64_t at = INT64_MAX;
  for (int64_t c = start + threadIdx.x; c < end; c += BLOCK) {
    double J[27];
    cell_closed_jd(u, ih, d, c, J);
#pragma unroll
    for (int p = 0; p < 27; ++p) {
      const double v = (J[p] - J[p] == 0.0) ? J[p] : __builtin_inf();
      if (v < best) {
        best = v;
        at = (c - start) * 27 + p;
      }
    }
  }
  __shared__ double sv[BLOCK];
  __shared__ int64_t si[BLOCK];
  sv[threadIdx.x] = best;
  si[threadIdx.x] = at;
  __syncthreads();
  for (int w = BLOCK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      const double ov = sv[threadIdx.x + w];
      const int64_t oi = si[threadIdx.x + w];
      if (ov < sv[threadIdx.x] || (ov == sv[threadIdx.x] && oi < si[threadIdx.x])) {
        sv[threadIdx.x] = ov;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  double* o = out + t * 26;
  if (threadIdx.x == 0) {
    const int64_t first = si[0] == INT64_MAX ? 0 : si[0];
    o[0] = sv[0];
    o[1] = (double)first;
  }
  const int64_t cell = start + (si[0] == INT64_MAX ? 0 : si[0]) / 27;
  if (threadIdx.x < 24) {
    const int a = threadIdx.x / 3, r = threadIdx.x % 3;
    o[2 + threadIdx.x] = u[3 * vertex_node(cell_base_node(cell, d), a, d) + r];
  }
}

// det(A + s V) - delta = c0 + c1 s + c2 s^2 + c3 s^3 at reference points of cells (jd_step.py:13-17): A = I +
// grad u and V = grad v at the point through the Q1 weights of the point (jd_step.py:204-222); c1 and c2 are the
// cofactor contractions tr(adj(A) V) and tr(A adj(V)). out: 4 doubles per point.
__global__ void __launch_bounds__(BLOCK) witness_coefficients_kernel(const double* __restrict__ u,
                                                                     const double* __restrict__ v,
                                                                     const double* __restrict__ k, Dims d,
                                                                     const int64_t* __restrict__ cells,
                                                                     const double* __restrict__ pts, int64_t n,
                                                                     double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  double ih[3];
  load_ih(k, ih);
  const double delta = k[K_DELTA];
  if (cells[i] < 0 || cells[i] >= d.ncell) {   // never read outside the field
    for (int q = 0; q < 4; ++q) out[4 * i + q] = __builtin_nan("");
    return;
  }
  const int64_t b = cell_base_node(cells[i], d);
  const double t[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
  double A[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}}, V[3][3] = {};
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int64_t node = vertex_node(b, a, d);
    double f[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) f[j] = (a >> j) & 1 ? t[j] : 1.0 - t[j];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double w = (((a >> j) & 1) ? ih[j] : -ih[j]) * f[other1(j)] * f[other2(j)];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        A[r][j] += w * u[3 * node + r];
        V[r][j] += w * v[3 * node + r];
      }
    }
  }
  double cA[3][3], cV[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      cA[r][j] = A[r1][j1] * A[r2][j2] - A[r1][j2] * A[r2][j1];
      cV[r][j] = V[r1][j1] * V[r2][j2] - V[r1][j2] * V[r2][j1];
    }
  double c1 = 0.0, c2 = 0.0;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      c1 += cA[r][j] * V[r][j];
      c2 += A[r][j] * cV[r][j];
    }
  double* o = out + 4 * i;
  o[0] = (A[0][0] * cA[0][0] + A[0][1] * cA[0][1] + A[0][2] * cA[0][2]) - delta;
  o[1] = c1;
  o[2] = c2;
  o[3] = V[0][0] * cV[0][0] + V[0][1] * cV[0][1] + V[0][2] * cV[0][2];
}

// mu of each row of n closed27 rows (jd_affine.py:130-151).
__global__ void __launch_bounds__(EXT_BLOCK) face_interior_rows_kernel(const double* __restrict__ J27, int64_t n,
                                                                       double* __restrict__ mu) {
  const int64_t i = (int64_t)blockIdx.x * EXT_BLOCK + threadIdx.x;
  if (i >= n) return;
  double J[27];
#pragma unroll
  for (int p = 0; p < 27; ++p) J[p] = J27[27 * i + p];
  mu[i] = t69::face_interior_min(J);
}

__global__ void __launch_bounds__(BLOCK) node_min_kernel(const double* __restrict__ minima, Dims d,
                                                         double* __restrict__ node) {
  const int64_t nnode = d.nx * d.ny * d.nz;
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= nnode) return;
  const int64_t x = i % d.nx, y = (i / d.nx) % d.ny, z = i / (d.nx * d.ny);
  double m = __builtin_inf();
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int64_t xa = x - (a & 1), ya = y - ((a >> 1) & 1), za = z - ((a >> 2) & 1);
    if (xa < 0 || xa >= d.cx || ya < 0 || ya >= d.cy || za < 0 || za >= d.cz) continue;
    m = fmin(m, minima[(za * d.cy + ya) * d.cx + xa]);
  }
  node[i] = m;
}

// ------------------------------------------------------------------------------------------------
// T66 first root in (0, 1] of b0 + b1 s + b2 s^2 + b3 s^3 with b0 > 0 (global_jd_step.py:22-87).
__device__ double first_root(double b0, double b1, double b2, double b3) {
  const double eps = 2.220446049250313e-16;
  const double nanv = __builtin_nan("");
  // critical points of 3 b3 s^2 + 2 b2 s + b1 in (0, 1)
  const double a = 3.0 * b3, bb = 2.0 * b2, cc = b1;
  const bool linear = a == 0.0;
  const double r_lin = (linear && bb != 0.0) ? -cc / bb : nanv;
  const double disc = bb * bb - 4.0 * a * cc;
  const double sq = sqrt(disc >= 0.0 ? disc : 0.0);
  const double q = -(bb + (bb >= 0.0 ? sq : -sq)) / 2.0;
  double r1 = (!linear && disc >= 0.0 && a != 0.0) ? q / a : nanv;
  double r2 = (!linear && disc >= 0.0 && q != 0.0) ? cc / q : nanv;
  r1 = linear ? r_lin : r1;
  r2 = linear ? nanv : r2;
  r1 = (r1 > 0.0 && r1 < 1.0) ? r1 : 1.0;
  r2 = (r2 > 0.0 && r2 < 1.0) ? r2 : 1.0;
  const double t1 = fmin(r1, r2), t2 = fmax(r1, r2);
  auto qv = [&](double s) { return b0 + s * (b1 + s * (b2 + s * b3)); };
  auto touch = [&](double s) {
    return 4.0 * eps * (fabs(b0) + s * (fabs(b1) + s * (fabs(b2) + s * fabs(b3))));
  };
  const bool q1 = qv(t1) <= touch(t1), q2 = qv(t2) <= touch(t2), q3 = qv(1.0) <= 0.0;
  double hi, lo;
  if (q1) { hi = t1; lo = 0.0; }
  else if (q2) { hi = t2; lo = t1; }
  else if (q3) { hi = 1.0; lo = t2; }
  else return __builtin_inf();
  for (int it = 0; it < 50; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (qv(mid) <= 0.0) hi = mid; else lo = mid;
  }
  return lo;
}


// The T66 cubic det(F + s V) - delta = c0 + c1 s + c2 s^2 + c3 s^3 at one point (F = I + grad u, V = grad v),
// scaled to s/L: b_i = c_i L^i (global_jd_step.py:90-129).
__device__ __forceinline__ void root_cubic(const double F[3][3], const double V[3][3], double delta, double L,
                                           double L2, double L3, double& c0, double& b1, double& b2, double& b3) {
  double AF[3][3], AV[3][3];
  cofactor_plain(F, AF);
  cofactor_plain(V, AV);
  c0 = ((F[0][0] * AF[0][0] + F[0][1] * AF[0][1]) + F[0][2] * AF[0][2]) - delta;
  double c1 = 0.0, c2 = 0.0;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      c1 += AF[r][j] * V[r][j];
      c2 += F[r][j] * AV[r][j];
    }
  const double c3 = (V[0][0] * AV[0][0] + V[0][1] * AV[0][1]) + V[0][2] * AV[0][2];
  b1 = c1 * L;
  b2 = c2 * L2;
  b3 = c3 * L3;
}

// Whether a cell's 27 T66 cubics can be skipped, from its edges alone: with Jlb and NF from cell_j_lower_bound for
// u, and NV = |w| >= |V(p)| (w_rj = max_k |e'_rjk|, the direction's edges), the scaled coefficients obey |b1| <=
// L NF^2/sqrt3 NV, |b2| <= L^2 NF NV^2/sqrt3, |b3| <= L^3 NV^3/(3 sqrt3). When Jlb - delta exceeds their sum, no
// point has c0 <= 0 and none a root in (0, 1]: the cubics would leave the sweep's results as they are. The margins
// (1e-9 relative on the sum, 1e-12 (1 + NF^3) absolute) cover the rounding of these bounds and of the cubics' own
// evaluation; a nonfinite edge (u's through the bound, v's through the 0 * e probe) keeps the cell.
__device__ __forceinline__ bool root_free_cell(double jlb, double NF, const Edges& Ev, double delta, double L) {
  double nv2 = 0.0, probe = 0.0;
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      double w = 0.0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        w = fmax(w, fabs(Ev.e[j][r][k]));
        probe = fma(Ev.e[j][r][k], 0.0, probe);
      }
      nv2 = fma(w, w, nv2);
    }
  const double inv_sqrt3 = 0.57735026918962584, inv_3sqrt3 = 0.19245008972987526;
  const double nv = sqrt(nv2);
  const double bound = L * (NF * NF * inv_sqrt3 * nv + L * (NF * nv2 * inv_sqrt3 + L * nv2 * nv * inv_3sqrt3));
  return probe == 0.0 && (jlb - delta) > bound * (1.0 + 1e-9) + 1e-12 * (1.0 + NF * NF * NF);
}

template <bool FREE>
__device__ __forceinline__ void load_direction(const double* __restrict__ v, int64_t c, int64_t b, const Dims& d,
                                               double w[8][3]) {
  if (FREE)
    load_cell_free(v, c, d, w);
  else
    load_cell(v, b, d, w);
}

// The T66 sweep in two launches. The screen (one lane per cell, few registers, three waves per SIMD) lists every
// wave with a cell that root_free_cell cannot clear; the cubic pass then evaluates only the listed waves. A cleared
// cell has no root, no c0 <= 0 point and no nonfinite coefficient, so the listed waves alone give the sweep's
// results; min, integer counts (exact in double) and max do not depend on the list's order.
// FREE: v is the direction's free-node vector (the node values of inject(v), read without the node field).
template <bool FREE>
__global__ void __launch_bounds__(BLOCK, 3) jd_root_screen_kernel(const double* __restrict__ u,
                                                                  const double* __restrict__ v,
                                                                  const double* __restrict__ k, double L_arg,
                                                                  const double* __restrict__ L_dev, Dims d,
                                                                  unsigned long long* __restrict__ count,
                                                                  int64_t* __restrict__ list) {
  const double L = L_dev ? *L_dev : L_arg;
  const int64_t c = xcd_block(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
  bool skip = true;
  if (c < d.ncell) {
    double ih[3];
    load_ih(k, ih);
    const int64_t b = cell_base_node(c, d);
    double jlb, NF;
    {
      Edges Eu;
      double w[8][3];
      load_cell(u, b, d, w);
      make_edges(w, ih, Eu);
      jlb = cell_j_lower_bound(Eu, NF);
    }
    Edges Ev;
    {
      double w[8][3];
      load_direction<FREE>(v, c, b, d, w);
      make_edges(w, ih, Ev);
    }
    skip = root_free_cell(jlb, NF, Ev, k[K_DELTA], L);
  }
  if (!__all(skip) && (threadIdx.x & 63) == 0) list[atomicAdd(count, 1ull)] = c >> 6;
}

// The cubic pass over the listed waves (wave w: cells 64 w .. 64 w + 63), each wave of the grid taking list entries
// with a stride; per-block [min root, count c0 <= 0, any nonfinite] (global_jd_step.py:90-129;
// torch_ops.py:636-673).
template <bool FREE>
__global__ void __launch_bounds__(BLOCK) jd_root_kernel(const double* __restrict__ u, const double* __restrict__ v,
                                                        const double* __restrict__ k, double L_arg,
                                                        const double* __restrict__ L_dev, Dims d,
                                                        const unsigned long long* __restrict__ count,
                                                        const int64_t* __restrict__ list,
                                                        double* __restrict__ partials) {
  const double L = L_dev ? *L_dev : L_arg;
  double best = __builtin_inf();
  double invalid = 0.0, bad = 0.0;
  const int64_t listed = (int64_t)*count;
  const int64_t waves = (int64_t)gridDim.x * (BLOCK / 64);
  for (int64_t i = (int64_t)blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6); i < listed; i += waves) {
    const int64_t c = list[i] * 64 + (threadIdx.x & 63);
    if (c >= d.ncell) continue;
    double ih[3];
    load_ih(k, ih);
    const double delta = k[K_DELTA];
    const double L2 = L * L, L3 = L * L * L;
    const int64_t b = cell_base_node(c, d);
    Edges Eu, Ev;
    {
      double w[8][3];
      load_cell(u, b, d, w);
      make_edges(w, ih, Eu);
      load_direction<FREE>(v, c, b, d, w);
      make_edges(w, ih, Ev);
    }
    // Pass 1: every point with its compile-time closed blend; nonfinite and c0 <= 0 points are counted, and a
    // point whose cubic may have a root in (0, 1] marks the lane.
    bool need = false;
    static_for<27>([&](auto P) {
      if constexpr (P % 9 == 0) {
        opaque_edges(Eu);
        opaque_edges(Ev);
      }
      double F[3][3], V[3][3];
      closed_point_grad<P>(Eu, F);
      F[0][0] += 1.0;
      F[1][1] += 1.0;
      F[2][2] += 1.0;
      closed_point_grad<P>(Ev, V);
      double c0, b1, b2, b3;
      root_cubic(F, V, delta, L, L2, L3, c0, b1, b2, b3);
      if (!(isfinite(c0) && isfinite(b1) && isfinite(b2) && isfinite(b3)))
        bad = 1.0;
      else if (!(c0 > 0.0))
        invalid += 1.0;
      else
        need |= !(c0 > fabs(b1) + fabs(b2) + fabs(b3));
    });
    // Pass 2, for a wave with a marked lane: the roots, over the points in order with the tabulated blends (the
    // same coefficients up to the sign of an exact zero, so the same points search).
    if (__any(need)) {
      // points p = 3 row + ix in order; the three points of an x row share each field's x column
#pragma unroll 1
      for (int row = 0; row < 9; ++row) {
        const double(&wr)[4] = CLOSED_W.w[3 * row][0];
        double Ux[3], Vx[3];
        column_w(Eu, 0, wr, Ux);
        column_w(Ev, 0, wr, Vx);
#pragma unroll 1
        for (int p = 3 * row; p < 3 * row + 3; ++p) {
          double F[3][3], V[3][3];
          point_grad_yz_w(Eu, CLOSED_W.w[p], Ux, F);
          F[0][0] += 1.0;
          F[1][1] += 1.0;
          F[2][2] += 1.0;
          point_grad_yz_w(Ev, CLOSED_W.w[p], Vx, V);
          double c0, b1, b2, b3;
          root_cubic(F, V, delta, L, L2, L3, c0, b1, b2, b3);
          if (isfinite(c0) && isfinite(b1) && isfinite(b2) && isfinite(b3) && c0 > 0.0 &&
              !(c0 > fabs(b1) + fabs(b2) + fabs(b3)))
            best = fmin(best, first_root(c0, b1, b2, b3));
        }
      }
    }
  }
  // min over the block; counts summed (exact integers in double)
  __shared__ double sh[BLOCK / 64][3];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  best = wave_min(best);
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    invalid += __shfl_xor(invalid, m, 64);
    bad = fmax(bad, __shfl_xor(bad, m, 64));
  }
  if (lane == 0) {
    sh[wave][0] = best;
    sh[wave][1] = invalid;
    sh[wave][2] = bad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double b0 = sh[0][0], i0 = sh[0][1], f0 = sh[0][2];
    for (int w = 1; w < BLOCK / 64; ++w) {
      b0 = fmin(b0, sh[w][0]);
      i0 += sh[w][1];
      f0 = fmax(f0, sh[w][2]);
    }
    partials[blockIdx.x * 3 + 0] = b0;
    partials[blockIdx.x * 3 + 1] = i0;
    partials[blockIdx.x * 3 + 2] = f0;
  }
}

__global__ void root_count_zero_kernel(unsigned long long* __restrict__ count) { *count = 0ull; }

__global__ void __launch_bounds__(RB) jd_root_reduce_kernel(const double* __restrict__ partials, int64_t nblocks,
                                                               double* __restrict__ out) {
  double best = __builtin_inf(), invalid = 0.0, bad = 0.0;
  constexpr int BATCH = 8;   // loads of eight blocks in flight at once; the fold order is unchanged
  for (int64_t b0 = threadIdx.x; b0 < nblocks; b0 += BATCH * RB) {
    double v[BATCH][3];
#pragma unroll
    for (int i = 0; i < BATCH; ++i) {
      const int64_t b = b0 + i * RB < nblocks ? b0 + i * RB : nblocks - 1;
#pragma unroll
      for (int q = 0; q < 3; ++q) v[i][q] = partials[3 * b + q];
    }
#pragma unroll
    for (int i = 0; i < BATCH; ++i) {
      const bool in = b0 + i * RB < nblocks;   // selected, not branched (keeps the loads ahead)
      best = in ? fmin(best, v[i][0]) : best;
      invalid = in ? invalid + v[i][1] : invalid;
      bad = in ? fmax(bad, v[i][2]) : bad;
    }
  }
  __shared__ double sh[RB / 64][3];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  best = wave_min(best);
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    invalid += __shfl_xor(invalid, m, 64);
    bad = fmax(bad, __shfl_xor(bad, m, 64));
  }
  if (lane == 0) {
    sh[wave][0] = best;
    sh[wave][1] = invalid;
    sh[wave][2] = bad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double b0 = sh[0][0], i0 = sh[0][1], f0 = sh[0][2];
    for (int w = 1; w < RB / 64; ++w) {
      b0 = fmin(b0, sh[w][0]);
      i0 += sh[w][1];
      f0 = fmax(f0, sh[w][2]);
    }
    out[0] = b0;
    out[1] = i0;
    out[2] = f0;
  }
}

// ------------------------------------------------------------------------------------------------
// Node-term energies: image, proximity and their stable differences (residuals.py:82-94,189-195).
__global__ void __launch_bounds__(BLOCK) node_energy_kernel(const double* __restrict__ warped,
                                                            const double* __restrict__ fixed,
                                                            const double* __restrict__ dw,
                                                            const double* __restrict__ u,
                                                            const double* __restrict__ ref,
                                                            const double* __restrict__ bwarped,
                                                            const double* __restrict__ bu, double lam, double omega,
                                                            Dims d, double* __restrict__ partials) {
  const int64_t nnode = d.nx * d.ny * d.nz;
  DD s[4] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  double none[1] = {0.0};
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < nnode; i += stride) {
    const int64_t x = i % d.nx, y = (i / d.nx) % d.ny, z = i / (d.nx * d.ny);
    const double tx = (x == 0 || x == d.nx - 1) ? 0.5 : 1.0;
    const double ty = (y == 0 || y == d.ny - 1) ? 0.5 : 1.0;
    const double tz = (z == 0 || z == d.nz - 1) ? 0.5 : 1.0;
    const double q = omega * tz * ty * tx;
    const double qd = dw ? q * dw[i] : q;
    const double r = warped[i] - fixed[i];
    dd_add(s[0], 0.5 * (qd * r * r));
    double pr = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double e = u[3 * i + c] - ref[3 * i + c];
      pr += q * (e * e);
    }
    dd_add(s[1], 0.5 * lam * pr);
    if (bwarped) {
      const double sq = sqrt(qd);
      const double xi = sq * (bwarped[i] - fixed[i]), yi = sq * r;
      dd_add(s[2], 0.5 * ((yi - xi) * (yi + xi)));
      const double sl = sqrt(lam * q);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double xp = sl * (bu[3 * i + c] - ref[3 * i + c]), yp = sl * (u[3 * i + c] - ref[3 * i + c]);
        dd_add(s[3], 0.5 * ((yp - xp) * (yp + xp)));
      }
    }
  }
  block_reduce_store<4, 0>(s, none, partials + blockIdx.x, gridDim.x);
}

// The energy gradient on the free nodes in one pass: for interior node f (trapezoid weight 1, so q = omega) the
// total node term qd (W - F) b + lam q (u - ref), written as node_terms_kernel writes it, plus the eight cell
// contributions in the order gather_kernel adds them, stored straight into the component-major free
// vector g[c F + f] (what restrict_kernel would make of the node field). Saves the node-term and restriction passes
// over the whole node field.
__global__ void __launch_bounds__(BLOCK) gather_free_kernel(const double* __restrict__ warped,
                                                            const double* __restrict__ fixed,
                                                            const double* __restrict__ dw,
                                                            const double* __restrict__ bgrad,
                                                            const double* __restrict__ u,
                                                            const double* __restrict__ ref, double lam, double omega,
                                                            const double* __restrict__ loc, Dims d,
                                                            double* __restrict__ g) {
  const int64_t mx = d.nx - 2, my = d.ny - 2, mz = d.nz - 2, F = mx * my * mz;
  const int64_t f = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (f >= F) return;
  const int64_t x = f % mx + 1, y = (f / mx) % my + 1, z = f / (mx * my) + 1;
  const int64_t i = (z * d.ny + y) * d.nx + x;
  const double q = omega * 1.0 * 1.0 * 1.0;
  const double qd = dw ? q * dw[i] : q;
  const double lq = lam * q;
  const double r = qd * (warped[i] - fixed[i]);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double im = r * bgrad[3 * i + c];
    const double acc = im + lq * (u[3 * i + c] - ref[3 * i + c]);
    g[c * F + f] = locals_gather<false>(loc, c, x, y, z, d, acc);
  }
}

// The three outer gradients on the free nodes in one pass (the joint form): gd = image node term, gf = image +
// barrier contributions, ge = total node term + soft contributions + barrier contributions, each gathered in the
// order gather_kernel uses (soft before barrier for ge) and stored component-major.
__global__ void __launch_bounds__(BLOCK) joint_free_kernel(const double* __restrict__ warped,
                                                           const double* __restrict__ fixed,
                                                           const double* __restrict__ dw,
                                                           const double* __restrict__ bgrad,
                                                           const double* __restrict__ u,
                                                           const double* __restrict__ ref, double lam, double omega,
                                                           const double* __restrict__ soft,
                                                           const double* __restrict__ bar, Dims d,
                                                           double* __restrict__ gd, double* __restrict__ gf,
                                                           double* __restrict__ ge) {
  const int64_t mx = d.nx - 2, my = d.ny - 2, mz = d.nz - 2, F = mx * my * mz;
  const int64_t f = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (f >= F) return;
  const int64_t x = f % mx + 1, y = (f / mx) % my + 1, z = f / (mx * my) + 1;
  const int64_t i = (z * d.ny + y) * d.nx + x;
  const double q = omega * 1.0 * 1.0 * 1.0;
  const double qd = dw ? q * dw[i] : q;
  const double lq = lam * q;
  const double r = qd * (warped[i] - fixed[i]);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double im = r * bgrad[3 * i + c];
    const double outer = locals_gather<false>(bar, c, x, y, z, d, im);
    double total = im + lq * (u[3 * i + c] - ref[3 * i + c]);
    total = locals_gather<false>(soft, c, x, y, z, d, total);
    total = locals_gather<false>(bar, c, x, y, z, d, total);
    gd[c * F + f] = im;
    gf[c * F + f] = outer;
    ge[c * F + f] = total;
  }
}

// Node terms of the gradient and of the GGN diagonal (torch_ops.py:563-568, 784-791;
// linearization.py:150-156, 362): image = qd (W - F) b; total = image + lam q (u - ref);
// ggn = qd b^2 + lam q. Any output may be NULL.
// One lane per vector component of a node row: a workgroup covers BLOCK of the 3 nx components of row (z, y), so
// the vector fields are read and written with consecutive addresses and the row indices come from the block id
// (no per-lane division); the node scalars are shared by the three components of a node (cache hits).
__global__ void __launch_bounds__(BLOCK) node_terms_kernel(const double* __restrict__ warped,
                                                           const double* __restrict__ fixed,
                                                           const double* __restrict__ dw,
                                                           const double* __restrict__ bgrad,
                                                           const double* __restrict__ u,
                                                           const double* __restrict__ ref, double lam, double omega,
                                                           Dims d, int64_t blocks_per_row, double* __restrict__ image,
                                                           double* __restrict__ total, double* __restrict__ ggn) {
  const int64_t row = blockIdx.x / blocks_per_row;
  const int64_t t = (blockIdx.x - row * blocks_per_row) * BLOCK + threadIdx.x;
  if (t >= 3 * d.nx) return;
  const int64_t y = row % d.ny, z = row / d.ny;
  const int64_t x = t / 3;
  const int64_t i = row * d.nx + x, e = 3 * row * d.nx + t;
  const double tx = (x == 0 || x == d.nx - 1) ? 0.5 : 1.0;
  const double ty = (y == 0 || y == d.ny - 1) ? 0.5 : 1.0;
  const double tz = (z == 0 || z == d.nz - 1) ? 0.5 : 1.0;
  const double q = omega * tz * ty * tx;
  const double qd = dw ? q * dw[i] : q;
  const double lq = lam * q;
  const double r = qd * (warped[i] - fixed[i]);
  const double b = bgrad[e];
  const double im = r * b;
  if (image) image[e] = im;
  if (total) total[e] = im + lq * (u[e] - ref[e]);
  if (ggn) ggn[e] = qd * (b * b) + lq;
}

}  // namespace seqvp

// ================================================================================================
// C-ABI
using namespace seqvp;

static inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// The T66 sweep's workspace: per-block partials of the cubic pass, the listed-wave count, the list.
struct RootWorkspace {
  int64_t eval_blocks;       // the cubic pass's grid: at most two workgroups per CU, at most one wave per list entry
  size_t count_offset, list_offset, bytes;
};

static RootWorkspace root_workspace(const Dims& d) {
  RootWorkspace r;
  const int64_t waves = cdiv(d.ncell, 64);
  r.eval_blocks = std::max<int64_t>(1, std::min<int64_t>(cdiv(waves, BLOCK / 64), 512));
  r.count_offset = (size_t)r.eval_blocks * 3 * sizeof(double);
  r.list_offset = r.count_offset + 256;
  r.bytes = r.list_offset + (size_t)waves * sizeof(int64_t);
  return r;
}

static size_t q1_workspace(int op, int64_t nx, int64_t ny, int64_t nz) {
  const Dims d = make_dims(nx, ny, nz);
  const int64_t nblocks = cdiv(d.ncell, BLOCK);
  switch (op) {
    case 0: return (size_t)nblocks * 6 * sizeof(double);
    case 1: return (size_t)nblocks * 6 * sizeof(double);
    case 2: return (size_t)d.ncell * 24 * sizeof(double);
    case 3: return (size_t)d.ncell * 24 * sizeof(double);
    case 4: return (size_t)d.ncell * 48 * sizeof(double);
    case 5: return root_workspace(d).bytes;
    case 8: return (size_t)d.ncell * sizeof(double);
    case 9: return (size_t)1024 * 8 * sizeof(double);
    case 10: return trial_stage_offset(nblocks) + TRIAL_STAGE_DOUBLES * sizeof(double);   // partials, least J, stage
    default: return 0;
  }
}

size_t seqvp_q1_workspace_bytes(int op, int64_t nx, int64_t ny, int64_t nz) { return q1_workspace(op, nx, ny, nz); }

static bool grid_ok(int64_t nx, int64_t ny, int64_t nz) { return nx >= 2 && ny >= 2 && nz >= 2; }

extern "C" int seqvp_q1_soft_energy(const double* u, const double* ref, const double* base, const double* k,
                                    int64_t nx, int64_t ny, int64_t nz, double* out, void* ws, size_t ws_bytes,
                                    void* stream) {
  if (!u || !ref || !k || !out || !grid_ok(nx, ny, nz)) return SEQVP_EINVAL;
  if (!ws || ws_bytes < q1_workspace(0, nx, ny, nz)) return SEQVP_EWORKSPACE;
  const Dims d = make_dims(nx, ny, nz);
  const int64_t nb = cdiv(d.ncell, BLOCK);
  hipStream_t s = (hipStream_t)stream;
  double* part = (double*)ws;
  if (base)
    hipLaunchKernelGGL((soft_energy_kernel<true>), dim3(nb), dim3(BLOCK), 0, s, u, ref, base, k, d, part);
  else
    hipLaunchKernelGGL((soft_energy_kernel<false>), dim3(nb), dim3(BLOCK), 0, s, u, ref, ref, k, d, part);
  hipLaunchKernelGGL((reduce_partials_kernel<3, 0>), dim3(1), dim3(RB), 0, s, part, nb, out);
  return launch_status();
}

extern "C" int seqvp_q1_trial_energy_base_min(const double* u, const double* ref, const double* base,
                                              const double* base_min, int quiet, const double* k, int64_t nx,
                                              int64_t ny, int64_t nz, double* out, void* ws, size_t ws_bytes,
                                              void* stream) {
  if (!u || !ref || !k || !out || !grid_ok(nx, ny, nz)) return SEQVP_EINVAL;
  if (quiet && (!base || !base_min || base == ref)) return SEQVP_EINVAL;
  if (!ws || ws_bytes < q1_workspace(10, nx, ny, nz)) return SEQVP_EWORKSPACE;
  const Dims d = make_dims(nx, ny, nz);
  const int64_t nb = cdiv(d.ncell, BLOCK);
  hipStream_t s = (hipStream_t)stream;
  double* soft = (double*)ws;
  double* bar = soft + nb * 6;
  unsigned long long* umin = (unsigned long long*)(bar + nb * 6);
  const double* bm = base ? base_min : nullptr;
  hipLaunchKernelGGL(umin_init_kernel, dim3(1), dim3(1), 0, s, umin);
  if (base == ref)      // the base is the reference: its Gauss-point terms are the reference's
    hipLaunchKernelGGL((trial_energy_kernel<true, true>), dim3(nb), dim3(BLOCK), 0, s, u, ref, ref, k, d, soft, bar,
                       bm, umin);
  else if (quiet)       // a base with no active closed27 point (the caller's guarantee)
    hipLaunchKernelGGL((trial_energy_kernel<true, false, true>), dim3(nb), dim3(BLOCK), 0, s, u, ref, base, k, d,
                       soft, bar, bm, umin);
  else if (base)
    hipLaunchKernelGGL((trial_energy_kernel<true>), dim3(nb), dim3(BLOCK), 0, s, u, ref, base, k, d, soft, bar, bm,
                       umin);
  else
    hipLaunchKernelGGL((trial_energy_kernel<false>), dim3(nb), dim3(BLOCK), 0, s, u, ref, ref, k, d, soft, bar,
                       (const double*)nullptr, umin);
  double* stage = (double*)((char*)ws + trial_stage_offset(nb));
  hipLaunchKernelGGL(reduce_trial_stage_kernel, dim3(4, 2), dim3(RB / 4), 0, s, soft, bar, nb, stage);
  hipLaunchKernelGGL(reduce_trial_final_kernel, dim3(1), dim3(64), 0, s, stage, out);
  return launch_status();
}

extern "C" int seqvp_q1_trial_energy(const double* u, const double* ref, const double* base, const double* k,
                                     int64_t nx, int64_t ny, int64_t nz, double* out, void* ws, size_t ws_bytes,
                                     void* stream) {
  return seqvp_q1_trial_energy_base_min(u, ref, base, nullptr, 0, k, nx, ny, nz, out, ws, ws_bytes, stream);
}

extern "C" int seqvp_q1_trial_energy_batch(const SeqvpTrialBatchEntry* pairs, int npairs, int tile, int64_t grid,
                                           void* stream) {
  if (!pairs || npairs < 1 || npairs > TRIAL_BATCH_MAX || tile < 1 || tile > 64 || grid < 0) return SEQVP_EINVAL;
  TrialBatchTable t = {};
  t.npairs = npairs;
  t.tile = tile;
  int64_t items = 0;
  for (int i = 0; i < npairs; ++i) {
    const SeqvpTrialBatchEntry& e = pairs[i];
    if (!e.u || !e.ref || !e.base || !e.k || !e.out || !grid_ok(e.nx, e.ny, e.nz)) return SEQVP_EINVAL;
    if (e.nx > INT32_MAX || e.ny > INT32_MAX || e.nz > INT32_MAX) return SEQVP_EINVAL;
    if (!e.workspace || e.workspace_bytes < q1_workspace(10, e.nx, e.ny, e.nz)) return SEQVP_EWORKSPACE;
    const Dims d = make_dims(e.nx, e.ny, e.nz);
    const int64_t nb = cdiv(d.ncell, BLOCK);
    const int64_t ni = cdiv(nb, tile);
    TrialBatchPair& P = t.p[i];
    P.u = e.u;
    P.ref = e.ref;
    P.base = e.base;
    P.k = e.k;
    P.soft_partials = (double*)e.workspace;
    P.barrier_partials = P.soft_partials + nb * 6;
    P.out = e.out;
    P.nx = (int32_t)e.nx;
    P.ny = (int32_t)e.ny;
    P.nz = (int32_t)e.nz;
    P.item_begin = (int32_t)items;
    P.nitems = (int32_t)ni;
    P.nblocks = (int32_t)nb;
    items += cdiv(ni, 8) * 8;
    if (items > INT32_MAX || nb > INT32_MAX) return SEQVP_EINVAL;
  }
  t.nitems = (int32_t)items;
  // a persistent grid is a multiple of 8 (items w, w + grid, ... then share w's XCD); 0 = one workgroup per item
  const int64_t g = grid == 0 ? items : std::min<int64_t>(cdiv(grid, 8) * 8, items);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(trial_energy_batch_kernel, dim3(g), dim3(BLOCK), 0, s, t);
  hipLaunchKernelGGL(trial_energy_batch_reduce_kernel, dim3(npairs, 2), dim3(RB), 0, s, t);
  return launch_status();
}

// ---- allocation-free soft energy ---------------------------------------------------------------------------------
// The block partials of seqvp_q1_soft_energy_noalloc live in this fixed device array (one copy per GPU, held until
// the process ends). Host code never uses its name as an address: each device's address comes from
// hipGetSymbolAddress on that device, resolved on the entry point's first call there. Every block writes its whole
// partial slot before the reduction reads it, so nothing needs clearing between calls.
constexpr int64_t NOALLOC_PARTIAL_DOUBLES = 1 << 19;      // 4 MiB: 6 doubles per 256-cell block
static __device__ __align__(256) double scratch[NOALLOC_PARTIAL_DOUBLES];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex scratch_mutex;
double* scratch_address[NOALLOC_MAX_DEVICES] = {};

// The device `stream` runs on (the current device for the null stream) and that device's scratch address.
int noalloc_scratch(hipStream_t stream, double** out) {
  int current = 0;
  if (hipGetDevice(&current) != hipSuccess) return SEQVP_ELAUNCH;
  int device = current;
  if (stream && hipStreamGetDevice(stream, &device) != hipSuccess) return SEQVP_EINVAL;
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return SEQVP_EINVAL;
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_address[device]) {
    if (device != current && hipSetDevice(device) != hipSuccess) return SEQVP_ELAUNCH;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (device != current) hipSetDevice(current);
    if (e != hipSuccess || !p) return SEQVP_ELAUNCH;
    scratch_address[device] = (double*)p;
  }
  *out = scratch_address[device];
  return SEQVP_OK;
}
}  // namespace

extern "C" int seqvp_q1_soft_energy_noalloc(const double* u, const double* ref, const double* base, const double* k,
                                            int64_t nx, int64_t ny, int64_t nz, double* out, void* stream) {
  if (!u || !ref || !k || !out || !grid_ok(nx, ny, nz)) return SEQVP_EINVAL;
  const Dims d = make_dims(nx, ny, nz);
  const int64_t nb = cdiv(d.ncell, BLOCK);
  if (nb * 6 > NOALLOC_PARTIAL_DOUBLES) return SEQVP_EWORKSPACE;     // grid beyond the fixed scratch
  hipStream_t s = (hipStream_t)stream;
  double* part = nullptr;
  const int rc = noalloc_scratch(s, &part);
  if (rc != SEQVP_OK) return rc;
  if (base)
    hipLaunchKernelGGL((soft_energy_kernel<true>), dim3(nb), dim3(BLOCK), 0, s, u, ref, base, k, d, part);
  else
    hipLaunchKernelGGL((soft_energy_kernel<false>), dim3(nb), dim3(BLOCK), 0, s, u, ref, ref, k, d, part);
  hipLaunchKernelGGL((reduce_partials_kernel<3, 0>), dim3(1), dim3(RB), 0, s, part, nb, out);
  return launch_status();
}

extern "C" int seqvp_q1_barrier_energy(const double* u, const double* base, const double* k, int64_t nx, int64_t ny,
                                       int64_t nz, double* out, void* ws, size_t ws_bytes, void* stream) {
  if (!u || !k || !out || !grid_ok(nx, ny, nz)) return SEQVP_EINVAL;
  if (!ws || ws_bytes < q1_workspace(1, nx, ny, nz)) return SEQVP_EWORKSPACE;
  const Dims d = make_dims(nx, ny, nz);
  const int64_t nb = cdiv(d.ncell, BLOCK);
  hipStream_t s = (hipStream_t)stream;
  double* part = (double*)ws;
  if (base)
    hipLaunchKernelGGL((barrier_energy_kernel<true>), dim3(nb), dim3(BLOCK), 0, s, u, base, k, d, part);
  else
    hipLaunchKernelGGL((barrier_energy_kernel<false>), dim3(nb), dim3(BLOCK), 0, s, u, u, k, d, part);
  // partials [val hi, val lo, chg hi, chg lo, minJ, min base J] -> out [val, minJ, chg, min base J]
  hipLaunchKernelGGL((reduce_partials_kernel<2, 2, true>), dim3(1), dim3(RB), 0, s, part, nb, out);
  return launch_status();
}

extern "C" int seqvp_q1_gradient(const double* node_base, const double* u, const double* ref, int soft, int barrier,
                                 const double* k, int64_t nx, int64_t ny, int64_t nz, double* out, void* ws,
                                 size_t ws_bytes, void* stream) {
  if (!u || !k || !out || !grid_ok(nx, ny, nz) || (soft && !ref)) return SEQVP_EINVAL;
  if (!ws || ws_bytes < q1_workspace(2, nx, ny, nz)) return SEQVP_EWORKSPACE;
  const Dims d = make_dims(nx, ny, nz);
  const int64_t nb = cdiv(d.ncell, BLOCK);
  hipStream_t s = (hipStream_t)stream;
  double* loc = (double*)ws;
  if (!soft && !barrier) {
    hipLaunchKernelGGL(copy_or_zero_kernel, dim3(cdiv(3 * nx * ny * nz, BLOCK)), dim3(BLOCK), 0, s, node_base, out,
                       3 * nx * ny * nz);
    return launch_status();
  }
  if (soft && barrier)
    hipLaunchKernelGGL((grad_locals_kernel<0, true, true>), dim3(nb), dim3(BLOCK), 0, s, u, ref, k, d, loc, loc);
  else if (soft)
    hipLaunchKernelGGL((grad_locals_kernel<0, true, false>), dim3(nb), dim3(BLOCK), 0, s, u, ref, k, d, loc, loc);
  else
    hipLaunchKernelGGL((grad_locals_kernel<1, false, true>), dim3(nb), dim3(BLOCK), 0, s, u, u, k, d, loc, loc);
  hipLaunchKernelGGL((gather_kernel<false>), dim3(cdiv(nx * ny * nz, BLOCK)), dim3(BLOCK), 0, s, node_base, loc,
                     loc, d, out);
  return launch_status();
}

extern "C" int seqvp_q1_gradient_free(const double* warped, const double* fixed, const double* data_weight,
                                      const double* image_grad, const double* u, const double* ref, double lam,
                                      double hx, double hy, double hz, const double* k, int64_t nx, int64_t ny,
                                      int64_t nz, double* g, void* ws, size_t ws_bytes, void* stream) {
  if (!warped || !fixed || !image_grad || !u || !ref || !k || !g || !grid_ok(nx, ny, nz)) return SEQVP_EINVAL;
  if (!ws || ws_bytes < q1_workspace(2, nx, ny, nz)) return SEQVP_EWORKSPACE;
  const Dims d = make_dims(nx, ny, nz);
  const int64_t nb = cdiv(d.ncell, BLOCK);
  const int64_t F = (nx - 2) * (ny - 2) * (nz - 2);
  hipStream_t s = (hipStream_t)stream;
  double* loc = (double*)ws;
  hipLaunchKernelGGL((grad_locals_kernel<0, true, true>), dim3(nb), dim3(BLOCK), 0, s, u, ref, k, d, loc, loc);
  if (F > 0)
    hipLaunchKernelGGL(gather_free_kernel, dim3(cdiv(F, BLOCK)), dim3(BLOCK), 0, s, warped, fixed, data_weight,
                       image_grad, u, ref, lam, hx * hy * hz, loc, d, g);
  return launch_status();
}

extern "C" int seqvp_q1_joint(const double* outer_base, const double* total_base, const double* u, const double* ref,
                              const double* k, int64_t nx, int64_t ny, int64_t nz, double* outer, double* total,
                              void* ws, size_t ws_bytes, void* stream) {
  if (!u || !ref || !k || !outer || !total || !grid_ok(nx, ny, nz)) return SEQVP_EINVAL;
  if (!ws || ws_bytes < q1_workspace(4, nx, ny, nz)) return SEQVP_EWORKSPACE;
  const Dims d = make_dims(nx, ny, nz);
  const int64_t nb = cdiv(d.ncell, BLOCK);
  hipStream_t s = (hipStream_t)stream;
  double* soft = (double*)ws;
  double* bar = soft + 24 * d.ncell;
  hipLaunchKernelGGL((grad_locals_kernel<2, true, true>), dim3(nb), dim3(BLOCK), 0, s, u, ref, k, d, soft, bar);
  const int64_t ng = cdiv(nx * ny * nz, BLOCK);
  hipLaunchKernelGGL((gather_kernel<false>), dim3(ng), dim3(BLOCK), 0, s, outer_base, bar, bar, d, outer);
  hipLaunchKernelGGL((gather_kernel<true>), dim3(ng), dim3(BLOCK), 0, s, total_base, soft, bar, d, total);
  return launch_status();
}

extern "C" int seqvp_q1_joint_free(const double* warped, const double* fixed, const double* data_weight,
                                   const double* image_grad, const double* u, const double* ref, double lam, double hx,
                                   double hy, double hz, const double* k, int64_t nx, int64_t ny, int64_t nz,
                                   double* gd, double* gf, double* ge, void* ws, size_t ws_bytes, void* stream) {
  if (!warped || !fixed || !image_grad || !u || !ref || !k || !gd || !gf || !ge || !grid_ok(nx, ny, nz))
    return SEQVP_EINVAL;
  if (!ws || ws_bytes < q1_workspace(4, nx, ny, nz)) return SEQVP_EWORKSPACE;
  const Dims d = make_dims(nx, ny, nz);
  const int64_t nb = cdiv(d.ncell, BLOCK);
  const int64_t F = (nx - 2) * (ny - 2) * (nz - 2);
  hipStream_t s = (hipStream_t)stream;
  double* soft = (double*)ws;
  double* bar = soft + 24 * d.ncell;
  hipLaunchKernelGGL((grad_locals_kernel<2, true, true>), dim3(nb), dim3(BLOCK), 0, s, u, ref, k, d, soft, bar);
  if (F > 0)
    hipLaunchKernelGGL(joint_free_kernel, dim3(cdiv(F, BLOCK)), dim3(BLOCK), 0, s, warped, fixed, data_weight,
                       image_grad, u, ref, lam, hx * hy * hz, soft, bar, d, gd, gf, ge);
  return launch_status();
}

extern "C" int seqvp_q1_ggn_diagonal(const double* node_base, const double* u, const double* k, int64_t nx,
                                     int64_t ny, int64_t nz, double* out, void* ws, size_t ws_bytes, void* stream) {
  if (!u || !k || !out || !grid_ok(nx, ny, nz)) return SEQVP_EINVAL;
  if (!ws || ws_bytes < q1_workspace(3, nx, ny, nz)) return SEQVP_EWORKSPACE;
  const Dims d = make_dims(nx, ny, nz);
  const int64_t nb = cdiv(d.ncell, BLOCK);
  hipStream_t s = (hipStream_t)stream;
  double* loc = (double*)ws;
  hipLaunchKernelGGL(ggn_locals_kernel, dim3(nb), dim3(BLOCK), 0, s, u, k, d, loc);
  hipLaunchKernelGGL((gather_kernel<false>), dim3(cdiv(nx * ny * nz, BLOCK)), dim3(BLOCK), 0, s, node_base, loc,
                     loc, d, out);
  return launch_status();
}

extern "C" int seqvp_node_min_jd(const double* u, const double* k, int64_t nx, int64_t ny, int64_t nz, double* node,
                                 void* ws, size_t ws_bytes, void* stream) {
  if (!u || !k || !node || !grid_ok(nx, ny, nz)) return SEQVP_EINVAL;
  if (!ws || ws_bytes < q1_workspace(8, nx, ny, nz)) return SEQVP_EWORKSPACE;
  const Dims d = make_dims(nx, ny, nz);
  hipStream_t s = (hipStream_t)stream;
  double* minima = (double*)ws;
  hipLaunchKernelGGL(cell_min_jd_kernel<false>, dim3(cdiv(d.ncell, BLOCK)), dim3(BLOCK), 0, s, u, k, d, 0.0, minima,
                     (uint32_t*)nullptr, (unsigned long long*)nullptr);
  hipLaunchKernelGGL(node_min_kernel, dim3(cdiv(nx * ny * nz, BLOCK)), dim3(BLOCK), 0, s, minima, d, node);
  return launch_status();
}

// Workspace of the extended measure: counters (4 x u64) | minima (ncell doubles) | candidate list (ncell u32).
static size_t node_jd_ext_workspace(int64_t nx, int64_t ny, int64_t nz) {
  const Dims d = make_dims(nx, ny, nz);
  return 32 + (size_t)d.ncell * (sizeof(double) + sizeof(uint32_t));
}

extern "C" size_t seqvp_node_min_jd_extended_workspace_bytes(int64_t nx, int64_t ny, int64_t nz) {
  return node_jd_ext_workspace(nx, ny, nz);
}

extern "C" int seqvp_node_min_jd_extended(const double* u, const double* k, int64_t nx, int64_t ny, int64_t nz,
                                          double activation, double* node, double* stats, void* ws, size_t ws_bytes,
                                          void* stream) {
  if (!u || !k || !node || !grid_ok(nx, ny, nz) || !(activation > 0.0) || !(activation <= 1.0)) return SEQVP_EINVAL;
  const Dims d = make_dims(nx, ny, nz);
  if (d.ncell > 0xffffffffll) return SEQVP_EINVAL;
  if (!ws || ws_bytes < node_jd_ext_workspace(nx, ny, nz)) return SEQVP_EWORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  unsigned long long* counters = (unsigned long long*)ws;
  double* minima = (double*)((char*)ws + 32);
  uint32_t* list = (uint32_t*)(minima + d.ncell);
  hipLaunchKernelGGL(jd_extension_init_kernel, dim3(1), dim3(64), 0, s, counters);
  hipLaunchKernelGGL(cell_min_jd_kernel<true>, dim3(cdiv(d.ncell, BLOCK)), dim3(BLOCK), 0, s, u, k, d, activation,
                     minima, list, counters);
  // one lane per candidate; the grid strides over the count the previous kernel left in counters[0]
  const int64_t eblocks = cdiv(d.ncell, EXT_BLOCK) < 4096 ? cdiv(d.ncell, EXT_BLOCK) : 4096;
  hipLaunchKernelGGL(jd_extension_kernel, dim3(eblocks), dim3(EXT_BLOCK), 0, s, u, k, d, list, counters, minima);
  hipLaunchKernelGGL(node_min_kernel, dim3(cdiv(nx * ny * nz, BLOCK)), dim3(BLOCK), 0, s, minima, d, node);
  if (stats) hipLaunchKernelGGL(jd_extension_stats_kernel, dim3(1), dim3(64), 0, s, counters, stats);
  return launch_status();
}

extern "C" int seqvp_closed27_tile_minima(const double* u, const double* k, int64_t nx, int64_t ny, int64_t nz,
                                          int64_t tile, double* out, void* stream) {
  if (!u || !k || !out || !grid_ok(nx, ny, nz) || tile < 1) return SEQVP_EINVAL;
  const Dims d = make_dims(nx, ny, nz);
  hipLaunchKernelGGL(closed27_tile_minima_kernel, dim3(cdiv(d.ncell, tile)), dim3(BLOCK), 0, (hipStream_t)stream,
                     u, k, d, tile, out);
  return launch_status();
}

extern "C" int seqvp_witness_coefficients(const double* u, const double* v, const double* k, int64_t nx, int64_t ny,
                                          int64_t nz, const int64_t* cells, const double* points, int64_t n,
                                          double* out, void* stream) {
  if (!u || !v || !k || !grid_ok(nx, ny, nz) || n < 0 || (n > 0 && (!cells || !points || !out))) return SEQVP_EINVAL;
  if (n == 0) return SEQVP_OK;
  const Dims d = make_dims(nx, ny, nz);
  hipLaunchKernelGGL(witness_coefficients_kernel, dim3(cdiv(n, BLOCK)), dim3(BLOCK), 0, (hipStream_t)stream, u, v,
                     k, d, cells, points, n, out);
  return launch_status();
}

extern "C" int seqvp_face_interior_minima(const double* J27, int64_t n, double* mu, void* stream) {
  if (n < 0 || (n > 0 && (!J27 || !mu))) return SEQVP_EINVAL;
  if (n == 0) return SEQVP_OK;
  hipLaunchKernelGGL(face_interior_rows_kernel, dim3(cdiv(n, EXT_BLOCK)), dim3(EXT_BLOCK), 0, (hipStream_t)stream,
                     J27, n, mu);
  return launch_status();
}

extern "C" int seqvp_face_interior_minima_host(const double* J27, int64_t n, double* mu) {
  if (n < 0 || (n > 0 && (!J27 || !mu))) return SEQVP_EINVAL;
  for (int64_t i = 0; i < n; ++i) mu[i] = t69::face_interior_min(J27 + 27 * i);
  return SEQVP_OK;
}

template <bool FREE>
static int jd_root_launch(const double* u, const double* v, const double* k, double limit, const double* limit_dev,
                          const Dims& d, double* out, void* ws, hipStream_t s) {
  const RootWorkspace r = root_workspace(d);
  double* part = (double*)ws;
  unsigned long long* count = (unsigned long long*)((char*)ws + r.count_offset);
  int64_t* list = (int64_t*)((char*)ws + r.list_offset);
  hipLaunchKernelGGL(root_count_zero_kernel, dim3(1), dim3(1), 0, s, count);
  hipLaunchKernelGGL((jd_root_screen_kernel<FREE>), dim3(cdiv(d.ncell, BLOCK)), dim3(BLOCK), 0, s, u, v, k, limit,
                     limit_dev, d, count, list);
  hipLaunchKernelGGL((jd_root_kernel<FREE>), dim3(r.eval_blocks), dim3(BLOCK), 0, s, u, v, k, limit, limit_dev, d,
                     count, list, part);
  hipLaunchKernelGGL(jd_root_reduce_kernel, dim3(1), dim3(RB), 0, s, part, r.eval_blocks, out);
  return launch_status();
}

extern "C" int seqvp_jd_root(const double* u, const double* v, const double* k, double limit, int64_t nx, int64_t ny,
                             int64_t nz, double* out, void* ws, size_t ws_bytes, void* stream) {
  if (!u || !v || !k || !out || !grid_ok(nx, ny, nz) || !(limit > 0.0)) return SEQVP_EINVAL;
  if (!ws || ws_bytes < q1_workspace(5, nx, ny, nz)) return SEQVP_EWORKSPACE;
  return jd_root_launch<false>(u, v, k, limit, nullptr, make_dims(nx, ny, nz), out, ws, (hipStream_t)stream);
}

extern "C" int seqvp_jd_root_free(const double* u, const double* p, const double* k, double limit, int64_t nx,
                                  int64_t ny, int64_t nz, double* out, void* ws, size_t ws_bytes, void* stream) {
  if (!u || !p || !k || !out || !grid_ok(nx, ny, nz) || !(limit > 0.0)) return SEQVP_EINVAL;
  if (!ws || ws_bytes < q1_workspace(5, nx, ny, nz)) return SEQVP_EWORKSPACE;
  return jd_root_launch<true>(u, p, k, limit, nullptr, make_dims(nx, ny, nz), out, ws, (hipStream_t)stream);
}

extern "C" int seqvp_jd_root_free_dev(const double* u, const double* p, const double* k, const double* limit,
                                      int64_t nx, int64_t ny, int64_t nz, double* out, void* ws, size_t ws_bytes,
                                      void* stream) {
  if (!u || !p || !k || !limit || !out || !grid_ok(nx, ny, nz)) return SEQVP_EINVAL;
  if (!ws || ws_bytes < q1_workspace(5, nx, ny, nz)) return SEQVP_EWORKSPACE;
  return jd_root_launch<true>(u, p, k, 0.0, limit, make_dims(nx, ny, nz), out, ws, (hipStream_t)stream);
}

extern "C" int seqvp_node_energy(const double* warped, const double* fixed, const double* data_weight,
                                 const double* u, const double* ref, const double* base_warped, const double* base_u,
                                 double lam, double hx, double hy, double hz, int64_t nx, int64_t ny, int64_t nz,
                                 double* out, void* ws, size_t ws_bytes, void* stream) {
  if (!warped || !fixed || !u || !ref || !out || !grid_ok(nx, ny, nz)) return SEQVP_EINVAL;
  if ((base_warped == nullptr) != (base_u == nullptr)) return SEQVP_EINVAL;
  if (!ws || ws_bytes < q1_workspace(9, nx, ny, nz)) return SEQVP_EWORKSPACE;
  const Dims d = make_dims(nx, ny, nz);
  hipStream_t s = (hipStream_t)stream;
  const int64_t nb = 1024;   // grid-stride; 1024 partials of 8 doubles
  double* part = (double*)ws;
  hipLaunchKernelGGL(node_energy_kernel, dim3(nb), dim3(BLOCK), 0, s, warped, fixed, data_weight, u, ref, base_warped,
                     base_u, lam, hx * hy * hz, d, part);
  hipLaunchKernelGGL((reduce_partials_kernel<4, 0>), dim3(1), dim3(RB), 0, s, part, nb, out);
  return launch_status();
}

extern "C" int seqvp_node_terms(const double* warped, const double* fixed, const double* data_weight,
                                const double* image_grad, const double* u, const double* ref, double lam, double hx,
                                double hy, double hz, int64_t nx, int64_t ny, int64_t nz, double* image, double* total,
                                double* ggn, void* stream) {
  if (!warped || !fixed || !image_grad || !grid_ok(nx, ny, nz)) return SEQVP_EINVAL;
  if (total && (!u || !ref)) return SEQVP_EINVAL;
  const Dims d = make_dims(nx, ny, nz);
  const int64_t per_row = cdiv(3 * nx, BLOCK);
  if (per_row * ny * nz > 0x7fffffff) return SEQVP_EINVAL;
  hipLaunchKernelGGL(node_terms_kernel, dim3(per_row * ny * nz), dim3(BLOCK), 0, (hipStream_t)stream, warped, fixed,
                     data_weight, image_grad, u, ref, lam, hx * hy * hz, d, per_row, image, total, ggn);
  return launch_status();
}

extern "C" int seqvp_q1_soft_energy_f32(const double* u, const double* ref, const double* base, const double* k,
                                        int64_t nx, int64_t ny, int64_t nz, double* out, void* ws, size_t ws_bytes,
                                        void* stream) {
  if (!u || !ref || !k || !out || !grid_ok(nx, ny, nz)) return SEQVP_EINVAL;
  if (!ws || ws_bytes < q1_workspace(0, nx, ny, nz)) return SEQVP_EWORKSPACE;
  const Dims d = make_dims(nx, ny, nz);
  const int64_t nb = cdiv(d.ncell, BLOCK);
  hipStream_t s = (hipStream_t)stream;
  double* part = (double*)ws;
  if (base)
    hipLaunchKernelGGL((soft_energy_f32_kernel<true>), dim3(nb), dim3(BLOCK), 0, s, u, ref, base, k, d, part);
  else
    hipLaunchKernelGGL((soft_energy_f32_kernel<false>), dim3(nb), dim3(BLOCK), 0, s, u, ref, ref, k, d, part);
  hipLaunchKernelGGL((reduce_partials_kernel<3, 0>), dim3(1), dim3(RB), 0, s, part, nb, out);
  return launch_status();
}

extern "C" int seqvp_q1_barrier_energy_f32(const double* u, const double* base, const double* k, int64_t nx,
                                           int64_t ny, int64_t nz, double* out, void* ws, size_t ws_bytes,
                                           void* stream) {
  if (!u || !k || !out || !grid_ok(nx, ny, nz)) return SEQVP_EINVAL;
  if (!ws || ws_bytes < q1_workspace(1, nx, ny, nz)) return SEQVP_EWORKSPACE;
  const Dims d = make_dims(nx, ny, nz);
  const int64_t nb = cdiv(d.ncell, BLOCK);
  hipStream_t s = (hipStream_t)stream;
  double* part = (double*)ws;
  if (base)
    hipLaunchKernelGGL((barrier_energy_f32_kernel<true>), dim3(nb), dim3(BLOCK), 0, s, u, base, k, d, part);
  else
    hipLaunchKernelGGL((barrier_energy_f32_kernel<false>), dim3(nb), dim3(BLOCK), 0, s, u, u, k, d, part);
  hipLaunchKernelGGL((reduce_partials_kernel<2, 2, true>), dim3(1), dim3(RB), 0, s, part, nb, out);
  return launch_status();
}
