// Whole-cell JD certificate of a candidate map: GPU screen (edge hull + T75 prescreen) and the
// authoritative CPU interval Bernstein fallback.
//
// Decision rule (jd_curl/three_d/feasibility.py:571-738, PrescreenCertificateConfig):
//   1. edge-hull margin per cell (compiled_enclosure.py:193-227); a 8192-cell tile holding any cell with a
//      nonfinite or inverted interval step loses its edge-hull proofs (feasibility.py:592-598);
//   2. cells the hull leaves unproven get the T75 prescreen; proven iff its margin > 0 (feasibility.py:663-695);
//   3. cells still unproven get the outward-rounded interval Bernstein fallback (jacobian_intervals,
//      lagrange_to_bernstein, _classify: feasibility.py:239-393, interval.py:106-124, compiled_fallback.py),
//      grouped in jobs of one tile's remaining cells, consumed in tile order (feasibility.py:697-738).
// Steps 1-2 run on the GPU over every cell (bitwise the CPU arithmetic: both share interval.h and
// prescreen_cell.h, compiled without contraction); the few undecided cells are gathered and classified on
// the CPU by worker threads. seqvp_certify_map_host runs all three steps on the CPU (the independent
// re-certification of an output).
#include <algorithm>
#include <atomic>
#include <queue>
#include <thread>
#include <vector>

#include "q1_common.h"
#include "interval.h"
#include "interval_simd.h"
#include "prescreen_cell.h"
#include "abi_util.h"

#pragma clang fp contract(off)

namespace seqvp {

using iv::I;

// ------------------------------------------------------------------------------------------------
// GPU screen
// ------------------------------------------------------------------------------------------------
static constexpr int CB = 256;

template <typename T>
__device__ __forceinline__ void load_vertices(const double* __restrict__ u, int64_t c, const Dims& d, T uv[8][3]) {
  const int64_t b = cell_base_node(c, d);
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const double* p = u + 3 * vertex_node(b, a, d);
    uv[a][0] = p[0];
    uv[a][1] = p[1];
    uv[a][2] = p[2];
  }
}

__global__ void __launch_bounds__(CB) zero_bytes_kernel(uint8_t* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * CB + threadIdx.x; i < n; i += (int64_t)gridDim.x * CB) p[i] = 0;
}

// margin[c] = lo(J_hull - delta); flag: the cell's own flag array (optional) and its tile's flag.
__global__ void __launch_bounds__(CB) edge_hull_kernel(const double* __restrict__ u, double hx, double hy, double hz,
                                                       double delta, Dims d, int64_t tile, double* __restrict__ margin,
                                                       uint8_t* __restrict__ cell_flag,
                                                       uint8_t* __restrict__ tile_flag) {
  const int64_t c = xcd_block(blockIdx.x, gridDim.x) * CB + threadIdx.x;
  if (c >= d.ncell) return;
  double uv[8][3];
  load_vertices(u, c, d, uv);
  const double h[3] = {hx, hy, hz};
  bool ok = true;
  margin[c] = iv::edge_hull_margin<true>(uv, h, delta, ok);
  if (cell_flag) cell_flag[c] = ok ? 0 : 1;
  if (!ok && tile_flag) tile_flag[c / tile] = 1;   // every writer stores the same byte
}

// Per cell: proven by the hull (code 0), by the prescreen (1) or undecided (2, appended to `list`).
// Block partials: min proven hull margin, min proven prescreen margin; counters: fast, prescreen, undecided.
__global__ void __launch_bounds__(CB) decide_kernel(const double* __restrict__ u, const double* __restrict__ D,
                                                    double delta, Dims d, int64_t tile,
                                                    const double* __restrict__ margin,
                                                    const uint8_t* __restrict__ tile_flag, uint8_t* __restrict__ codes,
                                                    int64_t* __restrict__ list, int64_t capacity,
                                                    double* __restrict__ part,
                                                    unsigned long long* __restrict__ counters) {
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t c = blk * CB + threadIdx.x;
  double mh = __builtin_inf(), mp = __builtin_inf();
  int code = 3;
  if (c < d.ncell) {
    const double m = margin[c];
    if (!tile_flag[c / tile] && m > 0.0) {
      code = 0;
      mh = m;
    }
  }
  // Waves whose cells the hull proved skip the prescreen entirely.
  if (__any(code == 3 && c < d.ncell)) {
    if (c < d.ncell && code == 3) {
      double v = __builtin_nan("");
      if (D) {
        double uv[8][3];
        load_vertices(u, c, d, uv);
        v = prescreen_vertices(uv, D);
      }
      const double pm = prescreen_margin(v, delta);
      if (pm > 0.0) {
        code = 1;
        mp = pm;
      } else {
        code = 2;
      }
    }
  }
  if (c < d.ncell && codes) codes[c] = (uint8_t)code;
  if (code == 2) {
    const unsigned long long slot = atomicAdd(counters + 2, 1ull);
    if ((int64_t)slot < capacity) list[slot] = c;
  }
  // block reductions (min is exact in any order; counts are integers)
  int nf = code == 0, np = code == 1;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    mh = fmin(mh, __shfl_xor(mh, o, 64));
    mp = fmin(mp, __shfl_xor(mp, o, 64));
    nf += __shfl_xor(nf, o, 64);
    np += __shfl_xor(np, o, 64);
  }
  __shared__ double smh[CB / 64], smp[CB / 64];
  __shared__ int snf[CB / 64], snp[CB / 64];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    smh[w] = mh;
    smp[w] = mp;
    snf[w] = nf;
    snp[w] = np;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = smh[0], b = smp[0];
    int f = snf[0], p = snp[0];
    for (int k = 1; k < CB / 64; ++k) {
      a = fmin(a, smh[k]);
      b = fmin(b, smp[k]);
      f += snf[k];
      p += snp[k];
    }
    // the counts go to the block's slots too (one atomic per block on a shared counter serialises 26k blocks)
    part[4 * blk] = a;
    part[4 * blk + 1] = b;
    part[4 * blk + 2] = (double)f;
    part[4 * blk + 3] = (double)p;
  }
}

// summary = [min hull margin, min prescreen margin, n_fast, n_prescreen, n_undecided] (+inf: none).
__global__ void __launch_bounds__(CB) screen_finish_kernel(const double* __restrict__ part, int64_t nblocks,
                                                           const unsigned long long* __restrict__ counters,
                                                           double* __restrict__ summary) {
  // minima exact in any order; the counts are integers below 2^53, so their double sums are exact too
  double a = __builtin_inf(), b = __builtin_inf(), f = 0.0, p = 0.0;
  constexpr int BATCH = 8;   // loads of eight blocks in flight at once; the fold order is unchanged
  for (int64_t i0 = threadIdx.x; i0 < nblocks; i0 += BATCH * CB) {
    double v[BATCH][4];
#pragma unroll
    for (int k = 0; k < BATCH; ++k) {
      const int64_t i = i0 + k * CB < nblocks ? i0 + k * CB : nblocks - 1;
#pragma unroll
      for (int q = 0; q < 4; ++q) v[k][q] = part[4 * i + q];
    }
#pragma unroll
    for (int k = 0; k < BATCH; ++k) {
      const bool in = i0 + k * CB < nblocks;   // selected, not branched (keeps the loads ahead)
      a = in ? fmin(a, v[k][0]) : a;
      b = in ? fmin(b, v[k][1]) : b;
      f = in ? f + v[k][2] : f;
      p = in ? p + v[k][3] : p;
    }
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    a = fmin(a, __shfl_xor(a, o, 64));
    b = fmin(b, __shfl_xor(b, o, 64));
    f += __shfl_xor(f, o, 64);
    p += __shfl_xor(p, o, 64);
  }
  __shared__ double sa[CB / 64], sb[CB / 64], sf[CB / 64], sp[CB / 64];
  if ((threadIdx.x & 63) == 0) {
    sa[threadIdx.x >> 6] = a;
    sb[threadIdx.x >> 6] = b;
    sf[threadIdx.x >> 6] = f;
    sp[threadIdx.x >> 6] = p;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < CB / 64; ++k) {
      a = fmin(a, sa[k]);
      b = fmin(b, sb[k]);
      f += sf[k];
      p += sp[k];
    }
    summary[0] = a;
    summary[1] = b;
    summary[2] = f;
    summary[3] = p;
    summary[4] = (double)counters[2];
  }
}

__global__ void __launch_bounds__(CB) zero_counters_kernel(unsigned long long* __restrict__ c) {
  if (threadIdx.x < 3) c[threadIdx.x] = 0;
}

// out[i][a][r] = u at vertex a of cells[i] (the (n, 8, 3) vertex array of a cell list).
__global__ void __launch_bounds__(CB) gather_vertices_kernel(const double* __restrict__ u,
                                                             const int64_t* __restrict__ cells, int64_t n, Dims d,
                                                             double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * CB + threadIdx.x;
  if (t >= n * 24) return;
  const int64_t i = t / 24;
  const int k = (int)(t % 24), a = k / 3, r = k % 3;
  const int64_t c = cells[i];
  if (c < 0 || c >= d.ncell) {
    out[t] = __builtin_nan("");
    return;
  }
  out[t] = u[3 * vertex_node(cell_base_node(c, d), a, d) + r];
}

// ------------------------------------------------------------------------------------------------
// CPU interval Bernstein fallback
// ------------------------------------------------------------------------------------------------
namespace host {

// Reference points {0,.5,1}^3, p = 9iz + 3iy + ix, coordinate k of point p.
static inline double point_coord(int p, int k) {
  const int i = k == 0 ? p % 3 : (k == 1 ? (p / 3) % 3 : p / 9);
  return 0.5 * i;
}

// The four j-edges (a, b = a | 1<<j) in vertex order and their exact dyadic bilinear weights at the 27
// points (compiled_fallback.py:36-57).
struct EdgeTables {
  int pair[3][4][2];
  double w[3][4][27];
  EdgeTables() {
    for (int j = 0; j < 3; ++j) {
      int e = 0;
      for (int a = 0; a < 8; ++a) {
        if (a & (1 << j)) continue;
        pair[j][e][0] = a;
        pair[j][e][1] = a | (1 << j);
        for (int p = 0; p < 27; ++p) {
          double weight = 1.0;
          for (int k = 0; k < 3; ++k) {
            if (k == j) continue;
            const double x = point_coord(p, k);
            weight *= ((a >> k) & 1) ? x : 1.0 - x;
          }
          w[j][e][p] = weight;
        }
        ++e;
      }
    }
  }
};
static const EdgeTables& edge_tables() {
  static const EdgeTables t;
  return t;
}

struct Arrays {
  I p[27];        // J - delta at the 27 points (Lagrange values)
  I c[27];        // Bernstein coefficients of J - delta
  double s[27];   // float64 sample J
};

// jacobian_intervals - delta, lagrange_to_bernstein and _sample_jacobians of one cell
// (feasibility.py:239-290, interval.py:106-113; operation order of compiled_fallback.py:62-160).
// Returns false when any interval step or sample is nonfinite or inverted.
static bool cell_arrays(const double v[8][3], const double h[3], double delta, Arrays& out) {
  const EdgeTables& T = edge_tables();
  bool ok = true;
  I F[27][3][3];
  for (int j = 0; j < 3; ++j) {
    I E[4][3];
    for (int e = 0; e < 4; ++e) {
      const int a = T.pair[j][e][0], b = T.pair[j][e][1];
      for (int r = 0; r < 3; ++r) {
        const double d = v[b][r] - v[a][r];
        const I s = iv::outward(d, d, ok);
        const double x1 = s.lo / h[j], x2 = s.hi / h[j];
        E[e][r] = iv::outward(iv::np_min(iv::np_min(iv::np_min(x1, x1), x2), x2),
                              iv::np_max(iv::np_max(iv::np_max(x1, x1), x2), x2), ok);
      }
    }
    for (int p = 0; p < 27; ++p)
      for (int r = 0; r < 3; ++r) {
        I acc{0.0, 0.0};
        for (int e = 0; e < 4; ++e) {
          const double w = T.w[j][e][p];
          const I t = iv::outward(E[e][r].lo * w, E[e][r].hi * w, ok);
          acc = iv::add(acc, t, ok);
        }
        F[p][r][j] = iv::add(acc, iv::point(r == j ? 1.0 : 0.0), ok);
      }
  }
  I b[27];
  for (int p = 0; p < 27; ++p) {
    const I J = iv::determinant(F[p], ok);
    out.p[p] = iv::sub(J, iv::point(delta), ok);
    b[p] = out.p[p];
  }
  // axes x, y, z in turn (storage index 9iz+3iy+ix): middle = 2b - (a + c) * .5
  const int strides[3] = {1, 3, 9};
  for (int pass = 0; pass < 3; ++pass) {
    const int st = strides[pass];
    for (int i0 = 0; i0 < 27; ++i0) {
      if ((i0 / st) % 3 != 0) continue;
      const int a0 = i0, a1 = i0 + st, a2 = i0 + 2 * st;
      const I m = iv::mul(b[a1], iv::point(2.0), ok);
      const I s = iv::add(b[a0], b[a2], ok);
      const I hh = iv::mul(s, iv::point(0.5), ok);
      b[a1] = iv::sub(m, hh, ok);
    }
  }
  for (int p = 0; p < 27; ++p) out.c[p] = b[p];
  // float64 samples
  double fs[27][3][3];
  for (int p = 0; p < 27; ++p)
    for (int r = 0; r < 3; ++r)
      for (int j = 0; j < 3; ++j) fs[p][r][j] = r == j ? 1.0 : 0.0;
  for (int j = 0; j < 3; ++j)
    for (int e = 0; e < 4; ++e) {
      const int a = T.pair[j][e][0], bb = T.pair[j][e][1];
      for (int p = 0; p < 27; ++p) {
        const double w = T.w[j][e][p];
        for (int r = 0; r < 3; ++r) fs[p][r][j] = fs[p][r][j] + ((v[bb][r] - v[a][r]) / h[j]) * w;
      }
    }
  for (int p = 0; p < 27; ++p) {
    const double(*f)[3] = fs[p];
    bool fin = true;
    for (int r = 0; r < 3; ++r)
      for (int j = 0; j < 3; ++j) fin = fin && iv::is_finite(f[r][j]);
    const double a00 = f[1][1] * f[2][2] - f[1][2] * f[2][1];
    const double a01 = f[1][2] * f[2][0] - f[1][0] * f[2][2];
    const double a02 = f[1][0] * f[2][1] - f[1][1] * f[2][0];
    const double a10 = f[2][1] * f[0][2] - f[2][2] * f[0][1];
    const double a11 = f[2][2] * f[0][0] - f[2][0] * f[0][2];
    const double a12 = f[2][0] * f[0][1] - f[2][1] * f[0][0];
    const double a20 = f[0][1] * f[1][2] - f[0][2] * f[1][1];
    const double a21 = f[0][2] * f[1][0] - f[0][0] * f[1][2];
    const double a22 = f[0][0] * f[1][1] - f[0][1] * f[1][0];
    const double cof[9] = {a00, a01, a02, a10, a11, a12, a20, a21, a22};
    for (double x : cof) fin = fin && iv::is_finite(x);
    const double t0 = f[0][0] * a00, t1 = f[0][1] * a01, t2 = f[0][2] * a02;
    double s = 0.0;
    s += t0;
    s += t1;
    s += t2;
    fin = fin && iv::is_finite(t0) && iv::is_finite(t1) && iv::is_finite(t2) && iv::is_finite(s);
    out.s[p] = s;
    ok = ok && fin;
  }
  return ok;
}

enum Status { CERTIFIED = 0, UNCERTAIN = 1, INFEASIBLE = 2 };
enum CellReason {
  R_POSITIVE_COVER = 0, R_SAMPLE_WITNESS = 1, R_SUBDIVISION_WITNESS = 2, R_BOX_BUDGET = 3, R_DEPTH_BUDGET = 4,
  R_ARITHMETIC = 5
};

struct CellResult {
  int status = UNCERTAIN;
  int reason = R_ARITHMETIC;
  bool has_margin = false;
  double margin = 0.0;      // min_margin_lower_bound
  double lower = 0.0;       // min_jd_lower_bound
  int64_t boxes = 0;
  int depth = 0;
  bool has_witness = false;
  double witness[3] = {0, 0, 0};
};

struct Box {
  I c[27];
  double low[3], high[3];
  int depth;
};

struct HeapItem {
  double key;
  int64_t serial;
  int box;     // index into the box store
};
struct HeapCmp {   // min-heap on (key, serial), the tuple order of heapq
  bool operator()(const HeapItem& a, const HeapItem& b) const {
    return a.key > b.key || (a.key == b.key && a.serial > b.serial);
  }
};

static inline double min_lo(const I c[27]) {
  double m = c[0].lo;
  for (int p = 1; p < 27; ++p) m = iv::np_min(m, c[p].lo);
  return m;
}

// split_bernstein at the midpoint along reference axis 0/1/2 (interval.py:116-124).
static bool split(const Box& cur, int axis, Box& left, Box& right) {
  bool ok = true;
  const int st = axis == 0 ? 1 : (axis == 1 ? 3 : 9);
  for (int i0 = 0; i0 < 27; ++i0) {
    if ((i0 / st) % 3 != 0) continue;
    const I a = cur.c[i0], b = cur.c[i0 + st], c = cur.c[i0 + 2 * st];
    const I ab = iv::mul(iv::add(a, b, ok), iv::point(0.5), ok);
    const I bc = iv::mul(iv::add(b, c, ok), iv::point(0.5), ok);
    const I mid = iv::mul(iv::add(ab, bc, ok), iv::point(0.5), ok);
    left.c[i0] = a;
    left.c[i0 + st] = ab;
    left.c[i0 + 2 * st] = mid;
    right.c[i0] = mid;
    right.c[i0 + st] = bc;
    right.c[i0 + 2 * st] = c;
  }
  return ok;
}

// _classify with the adaptive lowest-lower-bound-first policy (feasibility.py:305-372).
static CellResult classify(const Arrays& A, double delta, int max_depth, int64_t max_boxes) {
  const double root_lower = min_lo(A.c);
  auto result = [&](int status, double margin, int64_t boxes, int depth, const double* witness, int reason) {
    CellResult r;
    r.boxes = boxes;
    r.depth = depth;
    double lower = margin;
    if (delta != 0.0) {
      bool ok = true;
      lower = iv::add(iv::point(margin), iv::point(delta), ok).lo;
      if (!ok) {
        r.status = UNCERTAIN;
        r.reason = R_ARITHMETIC;
        return r;
      }
    }
    r.status = status;
    r.reason = reason;
    r.has_margin = true;
    r.margin = margin;
    r.lower = lower;
    if (witness) {
      r.has_witness = true;
      for (int k = 0; k < 3; ++k) r.witness[k] = witness[k];
    }
    return r;
  };
  for (int iz = 0; iz < 3; ++iz)
    for (int iy = 0; iy < 3; ++iy)
      for (int ix = 0; ix < 3; ++ix)
        if (A.p[9 * iz + 3 * iy + ix].hi <= 0.0) {
          const double w[3] = {ix / 2.0, iy / 2.0, iz / 2.0};
          return result(INFEASIBLE, root_lower, 1, 0, w, R_SAMPLE_WITNESS);
        }
  std::vector<Box> store;
  store.reserve(64);
  Box root;
  for (int p = 0; p < 27; ++p) root.c[p] = A.c[p];
  for (int k = 0; k < 3; ++k) {
    root.low[k] = 0.0;
    root.high[k] = 1.0;
  }
  root.depth = 0;
  store.push_back(root);
  std::priority_queue<HeapItem, std::vector<HeapItem>, HeapCmp> pending;
  pending.push(HeapItem{root_lower, 0, 0});
  int64_t serial = 0, boxes = 0;
  int depth_used = 0;
  bool uncertain = false, have_cert = false;
  double cert_min = 0.0;
  while (!pending.empty()) {
    if (boxes >= max_boxes) return result(UNCERTAIN, root_lower, boxes, depth_used, nullptr, R_BOX_BUDGET);
    const HeapItem top = pending.top();
    pending.pop();
    const Box cur = store[top.box];
    ++boxes;
    depth_used = std::max(depth_used, cur.depth);
    const double minimum = min_lo(cur.c);
    if (minimum > 0.0) {
      cert_min = have_cert ? std::min(cert_min, minimum) : minimum;
      have_cert = true;
      continue;
    }
    for (int kz = 0; kz < 2; ++kz)
      for (int ky = 0; ky < 2; ++ky)
        for (int kx = 0; kx < 2; ++kx)
          if (cur.c[9 * (2 * kz) + 3 * (2 * ky) + 2 * kx].hi <= 0.0) {
            const double w[3] = {kx ? cur.high[0] : cur.low[0], ky ? cur.high[1] : cur.low[1],
                                 kz ? cur.high[2] : cur.low[2]};
            return result(INFEASIBLE, root_lower, boxes, depth_used, w, R_SUBDIVISION_WITNESS);
          }
    if (cur.depth >= max_depth) {
      uncertain = true;
      continue;
    }
    int axis = 0;
    double best = cur.high[0] - cur.low[0];
    for (int k = 1; k < 3; ++k) {
      const double ext = cur.high[k] - cur.low[k];
      if (ext > best) {
        best = ext;
        axis = k;
      }
    }
    Box left, right;
    if (!split(cur, axis, left, right))
      return result(UNCERTAIN, root_lower, boxes, depth_used, nullptr, R_ARITHMETIC);
    const double midpoint = (cur.low[axis] + cur.high[axis]) * 0.5;
    for (int k = 0; k < 3; ++k) {
      left.low[k] = cur.low[k];
      left.high[k] = cur.high[k];
      right.low[k] = cur.low[k];
      right.high[k] = cur.high[k];
    }
    left.high[axis] = midpoint;
    right.low[axis] = midpoint;
    left.depth = right.depth = cur.depth + 1;
    store.push_back(left);
    pending.push(HeapItem{min_lo(left.c), ++serial, (int)store.size() - 1});
    store.push_back(right);
    pending.push(HeapItem{min_lo(right.c), ++serial, (int)store.size() - 1});
  }
  if (uncertain) return result(UNCERTAIN, root_lower, boxes, depth_used, nullptr, R_DEPTH_BUDGET);
  return result(CERTIFIED, cert_min, boxes, depth_used, nullptr, R_POSITIVE_COVER);
}

// One job: the remaining cells of one tile (feasibility.py:714-738).
struct JobResult {
  bool failed = false;          // arithmetic failure of the job's arrays (every cell uncertain)
  bool arith = false;           // some cell classified arithmetic_uncertain
  bool has_margin = false;
  double margin = 0.0;
  bool has_sample = false;
  double sampled = 0.0;
  int64_t count = 0, unknown = 0, bad = 0, boxes = 0;
  int depth = 0;
  int64_t witness_cell = -1;
  double witness[3] = {0, 0, 0};
};

static void take_margin(JobResult& j, double m) {
  j.margin = j.has_margin ? iv::np_min(j.margin, m) : m;
  j.has_margin = true;
}

static JobResult run_job(const double* vertices, const int64_t* cells, int64_t n, const double h[3], double delta,
                         int max_depth, int64_t max_boxes) {
  JobResult J;
  J.count = n;
  std::vector<Arrays> arrays((size_t)n);
  bool ok = true;
  for (int64_t i = 0; i < n; ++i) {
    const double(*v)[3] = reinterpret_cast<const double(*)[3]>(vertices + 24 * i);
    ok = cell_arrays(v, h, delta, arrays[(size_t)i]) && ok;
  }
  if (!ok) {
    J.failed = true;
    J.unknown = n;
    return J;
  }
  std::vector<uint8_t> safe((size_t)n);
  bool any_safe = false;
  double safe_min = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const Arrays& A = arrays[(size_t)i];
    for (int p = 0; p < 27; ++p) {
      J.sampled = J.has_sample ? iv::np_min(J.sampled, A.s[p]) : A.s[p];
      J.has_sample = true;
    }
    const double m = min_lo(A.c);
    safe[(size_t)i] = m > 0.0;
    if (m > 0.0) {
      safe_min = any_safe ? iv::np_min(safe_min, m) : m;
      any_safe = true;
    }
  }
  J.boxes += n;
  if (any_safe) take_margin(J, safe_min);
  for (int64_t i = 0; i < n; ++i) {
    if (safe[(size_t)i]) continue;
    const CellResult r = classify(arrays[(size_t)i], delta, max_depth, max_boxes);
    if (r.has_margin) take_margin(J, r.margin);
    if (r.reason == R_ARITHMETIC) J.arith = true;
    J.boxes += r.boxes - 1;
    J.depth = std::max(J.depth, r.depth);
    J.unknown += r.status == UNCERTAIN;
    J.bad += r.status == INFEASIBLE;
    if (r.status == INFEASIBLE && J.witness_cell < 0) {
      J.witness_cell = cells[i];
      for (int k = 0; k < 3; ++k) J.witness[k] = r.witness[k];
    }
  }
  return J;
}

// Runs fn(i) for i in [0, n) on up to `workers` threads (fn must be thread-safe per index).
template <typename Fn>
static void parallel_for(int64_t n, int workers, Fn fn) {
  workers = std::max(1, std::min<int>(workers, (int)std::min<int64_t>(n, 64)));
  if (workers <= 1) {
    for (int64_t i = 0; i < n; ++i) fn(i);
    return;
  }
  std::atomic<int64_t> next{0};
  std::vector<std::thread> pool;
  for (int w = 0; w < workers; ++w)
    pool.emplace_back([&]() {
      for (int64_t i; (i = next.fetch_add(1)) < n;) fn(i);
    });
  for (auto& t : pool) t.join();
}

// Aggregate of a whole certificate before the final status (feasibility.py:600-769).
struct Totals {
  int64_t checked = 0, unknown = 0, bad = 0, boxes = 0, fast = 0, prescreened = 0, fallback = 0;
  int depth = 0;
  bool arith = false, has_margin = false, has_sample = false;
  double margin = 0.0, sampled = 0.0;
  int64_t witness_cell = -1;
  double witness[3] = {0, 0, 0};
  void margin_in(double m) {
    margin = has_margin ? iv::np_min(margin, m) : m;
    has_margin = true;
  }
};

// Fallback over a sorted cell list grouped by tile (jobs in tile order), folded into `t`.
static void fallback(const double* vertices, const int64_t* cells, int64_t n, const double h[3], double delta,
                     int64_t tile, int max_depth, int64_t max_boxes, int workers, Totals& t) {
  std::vector<int64_t> starts;
  for (int64_t i = 0; i < n; ++i)
    if (i == 0 || cells[i] / tile != cells[i - 1] / tile) starts.push_back(i);
  starts.push_back(n);
  const int64_t njobs = (int64_t)starts.size() - 1;
  std::vector<JobResult> res((size_t)std::max<int64_t>(njobs, 0));
  parallel_for(njobs, workers, [&](int64_t k) {
    const int64_t s = starts[(size_t)k], e = starts[(size_t)k + 1];
    res[(size_t)k] = run_job(vertices + 24 * s, cells + s, e - s, h, delta, max_depth, max_boxes);
  });
  for (const JobResult& J : res) {
    if (J.failed) {
      t.arith = true;
      t.unknown += J.count;
      continue;
    }
    if (J.has_sample) {
      t.sampled = t.has_sample ? iv::np_min(t.sampled, J.sampled) : J.sampled;
      t.has_sample = true;
    }
    if (J.has_margin) t.margin_in(J.margin);
    t.arith = t.arith || J.arith;
    t.boxes += J.boxes;
    t.depth = std::max(t.depth, J.depth);
    t.unknown += J.unknown;
    t.bad += J.bad;
    if (J.witness_cell >= 0 && t.witness_cell < 0) {
      t.witness_cell = J.witness_cell;
      for (int k = 0; k < 3; ++k) t.witness[k] = J.witness[k];
    }
  }
}

static void finish(const Totals& t, double delta, SeqvpCertificate* out) {
  out->status = t.bad ? INFEASIBLE : (t.unknown ? UNCERTAIN : CERTIFIED);
  out->reason = t.bad ? SEQVP_CERT_JACOBIAN_VIOLATION
                      : (t.arith ? SEQVP_CERT_ARITHMETIC_UNCERTAIN
                                 : (t.unknown ? SEQVP_CERT_SUBDIVISION_BUDGET : SEQVP_CERT_POSITIVE_BERNSTEIN_COVER));
  out->arithmetic_failure = t.arith ? 1 : 0;
  out->max_depth = t.depth;
  out->n_cells_checked = t.checked;
  out->n_uncertain = t.unknown;
  out->n_infeasible = t.bad;
  out->n_boxes = t.boxes;
  out->n_fast_cells = t.fast;
  out->n_prescreen_cells = t.prescreened;
  out->n_fallback_cells = t.fallback;
  out->witness_cell = t.witness_cell;
  for (int k = 0; k < 3; ++k) out->witness_point[k] = t.witness_cell >= 0 ? t.witness[k] : 0.0;
  const double nan = __builtin_nan("");
  out->min_margin = nan;
  out->min_jd = nan;
  if (t.has_margin && !t.arith) {
    out->min_margin = t.margin;
    if (delta == 0.0) {
      out->min_jd = t.margin;
    } else {
      bool ok = true;
      const I s = iv::add(iv::point(t.margin), iv::point(delta), ok);
      out->min_jd = ok ? s.lo : nan;
    }
  }
  out->sampled_min_jd = t.has_sample ? t.sampled : nan;
  if (t.fast || t.prescreened) {
    out->sampled_min_jd = nan;   // incomplete samples are not a global sample minimum
    if (out->status == CERTIFIED) out->reason = SEQVP_CERT_POSITIVE_ENCLOSURE_COVER;
  }
}

// A result decided before any cell is examined: nothing checked, every cell counted uncertain
// (the initial record of certify_map, feasibility.py:408-409).
static void early(SeqvpCertificate* out, int status, int reason, int64_t ncell) {
  Totals t;
  finish(t, 0.0, out);
  out->status = status;
  out->reason = reason;
  out->n_cells_checked = 0;
  out->n_uncertain = ncell;
}

// Edge-hull margin and flag of cell c of u (x fastest) with the scalar operations.
static inline void hull_cell(const double* u, const Dims& d, int64_t nx, int64_t ny, int64_t c, const double h[3],
                             double delta, double* margin, uint8_t* flag) {
  const int64_t x0 = c % d.cx, y0 = (c / d.cx) % d.cy, z0 = c / (d.cx * d.cy);
  const int64_t b = (z0 * ny + y0) * nx + x0;
  double uv[8][3];
  for (int a = 0; a < 8; ++a) {
    const double* p = u + 3 * (b + (a & 1) + ((a >> 1) & 1) * nx + ((a >> 2) & 1) * nx * ny);
    uv[a][0] = p[0];
    uv[a][1] = p[1];
    uv[a][2] = p[2];
  }
  bool ok = true;
  margin[c] = iv::edge_hull_margin<true>(uv, h, delta, ok);   // unflagged margins: bitwise
  flag[c] = ok ? 0 : 1;
}

// Cells [c0, c1): four consecutive cells of a row per AVX2 evaluation (the same bits as hull_cell), the row ends
// one by one.
__attribute__((target("avx2"))) static void hull_range_avx2(const double* u, const Dims& d, int64_t nx, int64_t ny,
                                                            int64_t c0, int64_t c1, const double h[3],
                                                            double delta, double* margin, uint8_t* flag) {
  for (int64_t c = c0; c < c1;) {
    const int64_t x0 = c % d.cx;
    if (x0 + 4 > d.cx || c + 4 > c1) {
      hull_cell(u, d, nx, ny, c, h, delta, margin, flag);
      ++c;
      continue;
    }
    const int64_t y0 = (c / d.cx) % d.cy, z0 = c / (d.cx * d.cy);
    const int64_t b = (z0 * ny + y0) * nx + x0;
    iv4::d4 uv[8][3];
    for (int a = 0; a < 8; ++a) {
      const double* p = u + 3 * (b + (a & 1) + ((a >> 1) & 1) * nx + ((a >> 2) & 1) * nx * ny);
      for (int r = 0; r < 3; ++r) uv[a][r] = iv4::d4{p[r], p[3 + r], p[6 + r], p[9 + r]};
    }
    iv4::l4 ok = (iv4::l4)-1;
    const iv4::d4 m = iv4::edge_hull_margin(uv, h, delta, ok);
    for (int l = 0; l < 4; ++l) {
      margin[c + l] = m[l];
      flag[c + l] = ok[l] ? 0 : 1;
    }
    c += 4;
  }
}

// Edge-hull margins and flags of every cell, in chunks over `workers` threads.
static void hull_all(const double* u, const Dims& d, int64_t nx, int64_t ny, const double h[3], double delta,
                     int workers, double* margin, uint8_t* flag) {
  const int64_t chunk = 16384;
  const int64_t nchunks = (d.ncell + chunk - 1) / chunk;
  const bool simd = iv4::avx2_available();
  parallel_for(nchunks, workers, [&](int64_t k) {
    const int64_t c0 = k * chunk, c1 = std::min(d.ncell, (k + 1) * chunk);
    if (simd) {
      hull_range_avx2(u, d, nx, ny, c0, c1, h, delta, margin, flag);
    } else {
      for (int64_t c = c0; c < c1; ++c) hull_cell(u, d, nx, ny, c, h, delta, margin, flag);
    }
  });
}

}  // namespace host
}  // namespace seqvp

using namespace seqvp;

// ---------------------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------------------
extern "C" size_t seqvp_certificate_workspace_bytes(int64_t nx, int64_t ny, int64_t nz, int64_t tile) {
  if (nx < 2 || ny < 2 || nz < 2 || tile < 1) return 0;
  const Dims d = make_dims(nx, ny, nz);
  const int64_t nb = (d.ncell + CB - 1) / CB;
  const int64_t ntile = (d.ncell + tile - 1) / tile;
  return (size_t)(8 * d.ncell + 32 * nb + 32 + ((ntile + 255) / 256) * 256);
}

extern "C" int seqvp_edge_hull(const double* u, double hx, double hy, double hz, double delta, int64_t nx, int64_t ny,
                               int64_t nz, double* margin, uint8_t* flag, void* stream) {
  if (!u || !margin || nx < 2 || ny < 2 || nz < 2) return SEQVP_EINVAL;
  const Dims d = make_dims(nx, ny, nz);
  hipLaunchKernelGGL(edge_hull_kernel, dim3((d.ncell + CB - 1) / CB), dim3(CB), 0, (hipStream_t)stream, u, hx, hy, hz,
                     delta, d, d.ncell, margin, flag, (uint8_t*)nullptr);
  return launch_status();
}

extern "C" int seqvp_certificate_screen(const double* u, const double* D, double hx, double hy, double hz, double delta,
                                        int64_t nx, int64_t ny, int64_t nz, int64_t tile, uint8_t* codes,
                                        int64_t* undecided, int64_t capacity, double* summary, void* workspace,
                                        size_t workspace_bytes, void* stream) {
  if (!u || !summary || nx < 2 || ny < 2 || nz < 2 || tile < 1 || capacity < 0 || (capacity > 0 && !undecided))
    return SEQVP_EINVAL;
  if (!workspace || workspace_bytes < seqvp_certificate_workspace_bytes(nx, ny, nz, tile)) return SEQVP_EWORKSPACE;
  const Dims d = make_dims(nx, ny, nz);
  const int64_t nb = (d.ncell + CB - 1) / CB;
  const int64_t ntile = (d.ncell + tile - 1) / tile;
  char* ws = (char*)workspace;
  double* margin = (double*)ws;
  double* part = (double*)(ws + 8 * d.ncell);
  unsigned long long* counters = (unsigned long long*)(ws + 8 * d.ncell + 32 * nb);
  uint8_t* tile_flag = (uint8_t*)(ws + 8 * d.ncell + 32 * nb + 32);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(zero_bytes_kernel, dim3((unsigned)std::min<int64_t>((ntile + CB - 1) / CB, 1024)), dim3(CB), 0, s,
                     tile_flag, ntile);
  hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, counters);
  hipLaunchKernelGGL(edge_hull_kernel, dim3(nb), dim3(CB), 0, s, u, hx, hy, hz, delta, d, tile, margin,
                     (uint8_t*)nullptr, tile_flag);
  hipLaunchKernelGGL(decide_kernel, dim3(nb), dim3(CB), 0, s, u, D, delta, d, tile, margin, tile_flag, codes,
                     undecided, capacity, part, counters);
  hipLaunchKernelGGL(screen_finish_kernel, dim3(1), dim3(CB), 0, s, part, nb, counters, summary);
  return launch_status();
}

extern "C" int seqvp_gather_vertices(const double* u, const int64_t* cells, int64_t n, int64_t nx, int64_t ny,
                                     int64_t nz, double* out, void* stream) {
  if (!u || (n > 0 && (!cells || !out)) || n < 0 || nx < 2 || ny < 2 || nz < 2) return SEQVP_EINVAL;
  if (n == 0) return SEQVP_OK;
  const Dims d = make_dims(nx, ny, nz);
  hipLaunchKernelGGL(gather_vertices_kernel, dim3((n * 24 + CB - 1) / CB), dim3(CB), 0, (hipStream_t)stream, u, cells,
                     n, d, out);
  return launch_status();
}

extern "C" int seqvp_bernstein_arrays(const double* vertices, int64_t n, const double* spacing, double delta,
                                      double* p_lo, double* p_hi, double* c_lo, double* c_hi, double* samples) {
  if (n < 0 || (n > 0 && (!vertices || !spacing || !p_lo || !p_hi || !c_lo || !c_hi || !samples))) return SEQVP_EINVAL;
  int all_ok = 1;
  for (int64_t i = 0; i < n; ++i) {
    host::Arrays A;
    const double(*v)[3] = reinterpret_cast<const double(*)[3]>(vertices + 24 * i);
    if (!host::cell_arrays(v, spacing, delta, A)) all_ok = 0;
    for (int p = 0; p < 27; ++p) {
      p_lo[27 * i + p] = A.p[p].lo;
      p_hi[27 * i + p] = A.p[p].hi;
      c_lo[27 * i + p] = A.c[p].lo;
      c_hi[27 * i + p] = A.c[p].hi;
      samples[27 * i + p] = A.s[p];
    }
  }
  return all_ok;
}

extern "C" int seqvp_point_jd_interval_host(const double* vertices, const double* spacing, const double* point,
                                            double* out) {
  if (!vertices || !spacing || !point || !out) return SEQVP_EINVAL;
  for (int j = 0; j < 3; ++j)
    if (!(spacing[j] > 0.0) || !(point[j] >= 0.0) || !(point[j] <= 1.0)) return SEQVP_EINVAL;
  double v[8][3];
  for (int a = 0; a < 8; ++a)
    for (int r = 0; r < 3; ++r) v[a][r] = vertices[3 * a + r];
  bool ok = true;
  for (int a = 0; a < 8; ++a)
    for (int r = 0; r < 3; ++r) ok = ok && seqvp::iv::is_finite(v[a][r]);
  const seqvp::iv::I J = seqvp::iv::point_jd_interval(v, spacing, point, ok);
  out[0] = ok ? J.lo : __builtin_nan("");
  out[1] = ok ? J.hi : __builtin_nan("");
  return SEQVP_OK;
}

extern "C" int seqvp_bernstein_cell(const double* vertices, const double* spacing, double delta, int max_depth,
                                    int64_t max_boxes, SeqvpCellCertificate* out) {
  if (!vertices || !spacing || !out || max_depth < 0 || max_boxes < 1) return SEQVP_EINVAL;
  const double(*v)[3] = reinterpret_cast<const double(*)[3]>(vertices);
  for (int k = 0; k < 24; ++k)
    if (!iv::is_finite(vertices[k])) {
      *out = SeqvpCellCertificate{};
      out->status = host::UNCERTAIN;
      out->reason = host::R_ARITHMETIC;
      out->min_margin = out->min_jd = __builtin_nan("");
      return SEQVP_OK;
    }
  host::Arrays A;
  host::CellResult r;
  if (host::cell_arrays(v, spacing, delta, A)) r = host::classify(A, delta, max_depth, max_boxes);
  out->status = r.status;
  out->reason = r.reason;
  out->n_boxes = r.boxes;
  out->max_depth = r.depth;
  out->min_margin = r.has_margin ? r.margin : __builtin_nan("");
  out->min_jd = r.has_margin ? r.lower : __builtin_nan("");
  out->has_witness = r.has_witness ? 1 : 0;
  for (int k = 0; k < 3; ++k) out->witness_point[k] = r.witness[k];
  double sm = A.s[0];
  for (int p = 1; p < 27; ++p) sm = iv::np_min(sm, A.s[p]);
  out->sampled_min_jd = sm;
  return SEQVP_OK;
}

extern "C" int seqvp_certificate_finish(const double* screen, int64_t ncell, const double* vertices,
                                        const int64_t* cells, int64_t n, const double* spacing, double delta,
                                        int64_t tile, int max_depth, int64_t max_boxes, int workers,
                                        SeqvpCertificate* out) {
  if (!screen || !spacing || !out || n < 0 || (n > 0 && (!vertices || !cells)) || tile < 1 || max_depth < 0 ||
      max_boxes < 1)
    return SEQVP_EINVAL;
  for (int64_t i = 1; i < n; ++i)
    if (cells[i] <= cells[i - 1]) return SEQVP_EINVAL;   // strictly increasing (tile order)
  host::Totals t;
  t.checked = ncell;
  t.fast = (int64_t)screen[2];
  t.prescreened = (int64_t)screen[3];
  t.fallback = n;
  t.boxes = t.fast + t.prescreened;
  if (t.fast && screen[0] < __builtin_inf()) t.margin_in(screen[0]);
  if (t.prescreened && screen[1] < __builtin_inf()) t.margin_in(screen[1]);
  host::fallback(vertices, cells, n, spacing, delta, tile, max_depth, max_boxes, workers, t);
  host::finish(t, delta, out);
  return SEQVP_OK;
}

extern "C" int seqvp_edge_hull_host(const double* u, int64_t nx, int64_t ny, int64_t nz, const double* spacing,
                                    double delta, int workers, double* margin, uint8_t* flag) {
  if (!u || !spacing || !margin || !flag || nx < 2 || ny < 2 || nz < 2) return SEQVP_EINVAL;
  host::hull_all(u, make_dims(nx, ny, nz), nx, ny, spacing, delta, workers, margin, flag);
  return SEQVP_OK;
}

extern "C" int seqvp_certify_map_host(const double* u, int64_t nx, int64_t ny, int64_t nz, const double* ax,
                                      const double* ay, const double* az, const double* spacing, double delta,
                                      const double* D, int64_t tile, int max_depth, int64_t max_boxes, int workers,
                                      SeqvpCertificate* out) {
  if (!u || !ax || !ay || !az || !spacing || !out || nx < 2 || ny < 2 || nz < 2 || tile < 1 || max_depth < 0 ||
      max_boxes < 1)
    return SEQVP_EINVAL;
  const Dims d = make_dims(nx, ny, nz);
  // NONFINITE_STATE (status uncertain), NONZERO_BOUNDARY, sampling domain (feasibility.py:413-431). Each z plane is
  // checked for all three on its own thread; the first kind found anywhere, in that order, decides, as three
  // whole-map passes would.
  {
    const double* axes[3] = {ax, ay, az};
    const int64_t len[3] = {nx, ny, nz};
    std::vector<uint8_t> found((size_t)(3 * nz), 0);
    host::parallel_for(nz, workers, [&](int64_t z) {
      bool nonfinite = false, boundary = false, domain = false;
      for (int64_t y = 0; y < ny; ++y)
        for (int64_t x = 0; x < nx; ++x) {
          const int64_t idx[3] = {x, y, z};
          const bool face = z == 0 || z == nz - 1 || y == 0 || y == ny - 1 || x == 0 || x == nx - 1;
          const double* p = u + 3 * ((z * ny + y) * nx + x);
          for (int j = 0; j < 3; ++j) {
            if (!iv::is_finite(p[j])) nonfinite = true;
            if (face && p[j] != 0.0) boundary = true;
            const double* A = axes[j];
            const double v = p[j] + A[idx[j]];
            const double lo = A[0], hi = A[len[j] - 1];
            if (!iv::is_finite(v) || v < lo || v > hi || (!face && (v <= lo || v >= hi))) domain = true;
          }
        }
      found[(size_t)(3 * z)] = nonfinite;
      found[(size_t)(3 * z + 1)] = boundary;
      found[(size_t)(3 * z + 2)] = domain;
    });
    bool any[3] = {false, false, false};
    for (int64_t z = 0; z < nz; ++z)
      for (int k = 0; k < 3; ++k) any[k] = any[k] || found[(size_t)(3 * z + k)];
    if (any[0]) {
      host::early(out, host::UNCERTAIN, SEQVP_CERT_NONFINITE_STATE, d.ncell);
      return SEQVP_OK;
    }
    if (any[1]) {
      host::early(out, host::INFEASIBLE, SEQVP_CERT_NONZERO_BOUNDARY, d.ncell);
      return SEQVP_OK;
    }
    if (any[2]) {
      host::early(out, host::INFEASIBLE, SEQVP_CERT_INVALID_DOMAIN, d.ncell);
      return SEQVP_OK;
    }
  }
  // 1. edge-hull margins (parallel, per-cell independent) and the tile rule: a tile with a flagged cell leaves all
  // its cells to the later stages, otherwise its cells with a positive margin are proven. Per tile in parallel;
  // the minimum of the proven (finite, positive) margins does not depend on the order it is taken in.
  std::vector<double> margin((size_t)d.ncell);
  std::vector<uint8_t> flag((size_t)d.ncell);
  const int64_t chunk = 16384;
  host::hull_all(u, d, nx, ny, spacing, delta, workers, margin.data(), flag.data());
  host::Totals t;
  t.checked = d.ncell;
  const int64_t ntiles = (d.ncell + tile - 1) / tile;
  std::vector<int64_t> proven((size_t)ntiles, 0);
  std::vector<double> least((size_t)ntiles, 0.0);
  std::vector<uint8_t> clean((size_t)ntiles, 0);
  host::parallel_for(ntiles, workers, [&](int64_t k) {
    const int64_t s0 = k * tile, e = std::min(d.ncell, s0 + tile);
    bool flagged = false;
    for (int64_t c = s0; c < e; ++c) flagged = flagged || flag[(size_t)c];
    clean[(size_t)k] = !flagged;
    if (flagged) return;
    int64_t n = 0;
    double m = 0.0;
    for (int64_t c = s0; c < e; ++c)
      if (margin[(size_t)c] > 0.0) {
        m = n ? (margin[(size_t)c] < m ? margin[(size_t)c] : m) : margin[(size_t)c];
        ++n;
      }
    proven[(size_t)k] = n;
    least[(size_t)k] = m;
  });
  std::vector<int64_t> rest;
  for (int64_t k = 0; k < ntiles; ++k) {
    const int64_t s0 = k * tile, e = std::min(d.ncell, s0 + tile);
    if (proven[(size_t)k]) {
      t.margin_in(least[(size_t)k]);
      t.fast += proven[(size_t)k];
    }
    if (proven[(size_t)k] == e - s0) continue;
    for (int64_t c = s0; c < e; ++c)
      if (!(clean[(size_t)k] && margin[(size_t)c] > 0.0)) rest.push_back(c);
  }
  t.boxes = t.fast;
  // 2. prescreen of the cells the hull left (when a derivative table is given)
  if (D && !rest.empty()) {
    std::vector<double> pm(rest.size());
    const int64_t nr = (int64_t)rest.size();
    host::parallel_for((nr + chunk - 1) / chunk, workers, [&](int64_t k) {
      double uv[8][3];
      for (int64_t i = k * chunk; i < std::min(nr, (k + 1) * chunk); ++i) {
        const int64_t c = rest[(size_t)i];
        const int64_t x0 = c % d.cx, y0 = (c / d.cx) % d.cy, z0 = c / (d.cx * d.cy);
        const int64_t b = (z0 * ny + y0) * nx + x0;
        for (int a = 0; a < 8; ++a) {
          const double* p = u + 3 * (b + (a & 1) + ((a >> 1) & 1) * nx + ((a >> 2) & 1) * nx * ny);
          uv[a][0] = p[0];
          uv[a][1] = p[1];
          uv[a][2] = p[2];
        }
        pm[(size_t)i] = prescreen_margin(prescreen_vertices(uv, D), delta);
      }
    });
    std::vector<int64_t> kept;
    for (size_t i = 0; i < rest.size(); ++i) {
      if (pm[i] > 0.0) {
        t.margin_in(pm[i]);
        ++t.prescreened;
      } else {
        kept.push_back(rest[i]);
      }
    }
    t.boxes += t.prescreened;
    rest.swap(kept);
  }
  t.fallback = (int64_t)rest.size();
  // 3. interval Bernstein fallback of what remains
  std::vector<double> verts(rest.size() * 24);
  for (size_t i = 0; i < rest.size(); ++i) {
    const int64_t c = rest[i];
    const int64_t x0 = c % d.cx, y0 = (c / d.cx) % d.cy, z0 = c / (d.cx * d.cy);
    const int64_t b = (z0 * ny + y0) * nx + x0;
    for (int a = 0; a < 8; ++a) {
      const double* p = u + 3 * (b + (a & 1) + ((a >> 1) & 1) * nx + ((a >> 2) & 1) * nx * ny);
      for (int r = 0; r < 3; ++r) verts[24 * i + 3 * a + r] = p[r];
    }
  }
  host::fallback(verts.data(), rest.data(), (int64_t)rest.size(), spacing, delta, tile, max_depth, max_boxes, workers,
                 t);
  host::finish(t, delta, out);
  return SEQVP_OK;
}
