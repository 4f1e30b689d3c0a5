// Binary64 interval arithmetic with one outward step per elementary operation, shared by the GPU
// certificate screen and the CPU certificate (host and device compile the same code).
//
// Semantics follow jd_curl/three_d/interval.py:31-124 and the per-cell kernels of
// compiled_enclosure.py:143-227 / compiled_fallback.py:62-160: every +, -, *, / of two intervals takes the
// rounded-to-nearest endpoint results (for * and / the min and max of the four endpoint products) and
// widens them by one ulp outward (nextafter towards -inf / +inf); an operation is valid when both widened
// endpoints are finite and ordered. Units including this header are compiled without floating-point
// contraction, so each operation is the separately rounded IEEE result the reference computes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#pragma clang fp contract(off)

namespace seqvp {
namespace iv {

__host__ __device__ __forceinline__ uint64_t bits_of(double x) {
  uint64_t b;
  memcpy(&b, &x, 8);
  return b;
}
__host__ __device__ __forceinline__ double from_bits(uint64_t b) {
  double x;
  memcpy(&x, &b, 8);
  return x;
}
__host__ __device__ __forceinline__ bool is_nan(double x) { return !(x == x); }
__host__ __device__ __forceinline__ bool is_finite(double x) { return x - x == 0.0; }

// nextafter(x, +inf) on the bits: the successor of x in the ordered binary64 line; +0 and -0 both step to
// the smallest positive subnormal, +inf and NaN are fixed points.
__host__ __device__ __forceinline__ double next_up(double x) {
  if (is_nan(x) || x == __builtin_inf()) return x;
  if (x == 0.0) return from_bits(1ull);
  const uint64_t b = bits_of(x);
  return from_bits(x > 0.0 ? b + 1 : b - 1);
}
// nextafter(x, -inf).
__host__ __device__ __forceinline__ double next_down(double x) { return -next_up(-x); }

// Variants for finite operands: equal to next_up/next_down for every finite x; for a nonfinite x the result is
// nonfinite too (NaN or inf), so an operation's validity check fails exactly where the full version's does.
__host__ __device__ __forceinline__ double next_up_finite(double x) {
  if (x == 0.0) return from_bits(1ull);
  const uint64_t b = bits_of(x);
  return from_bits(x > 0.0 ? b + 1 : b - 1);
}
__host__ __device__ __forceinline__ double next_down_finite(double x) { return -next_up_finite(-x); }

// 1/h when h is a positive power of two 2^k whose reciprocal is a normal number (|k| < 1022), else 0. For such h the
// quotient x / h and the product x * 2^-k are roundings of the same real number, so they are the same double for
// every x (IEEE division and multiplication are both correctly rounded): dividing by the spacing then costs one
// multiply instead of a division sequence, with identical bits.
__host__ __device__ __forceinline__ double pow2_reciprocal(double h) {
  const uint64_t b = bits_of(h);
  const uint64_t e = (b >> 52) & 0x7ff;
  if ((b >> 63) || (b & 0x000fffffffffffffull) || e < 2 || e > 2044) return 0.0;
  return from_bits((uint64_t)(2046 - e) << 52);
}
__host__ __device__ __forceinline__ double div_by(double x, double h, double inv) { return inv != 0.0 ? x * inv : x / h; }

// NumPy minimum/maximum: NaN propagates; otherwise the comparison result (ties return the second operand).
__host__ __device__ __forceinline__ double np_min(double a, double b) {
  if (is_nan(a)) return a;
  if (is_nan(b)) return b;
  return a < b ? a : b;
}
__host__ __device__ __forceinline__ double np_max(double a, double b) {
  if (is_nan(a)) return a;
  if (is_nan(b)) return b;
  return a > b ? a : b;
}

struct I {
  double lo, hi;
};

// outward(lo, hi) and its validity (finite, ordered).
__host__ __device__ __forceinline__ I outward(double lo, double hi, bool& ok) {
  const I r{next_down(lo), next_up(hi)};
  ok = ok && is_finite(r.lo) && is_finite(r.hi) && r.lo <= r.hi;
  return r;
}
__host__ __device__ __forceinline__ I add(I a, I b, bool& ok) { return outward(a.lo + b.lo, a.hi + b.hi, ok); }
__host__ __device__ __forceinline__ I sub(I a, I b, bool& ok) { return outward(a.lo - b.hi, a.hi - b.lo, ok); }
__host__ __device__ __forceinline__ I mul(I a, I b, bool& ok) {
  const double x = a.lo * b.lo, y = a.lo * b.hi, z = a.hi * b.lo, w = a.hi * b.hi;
  return outward(np_min(np_min(np_min(x, y), z), w), np_max(np_max(np_max(x, y), z), w), ok);
}
__host__ __device__ __forceinline__ I point(double v) { return I{v, v}; }
// Division; a divisor interval containing zero is invalid (interval.py:87-94).
__host__ __device__ __forceinline__ I div(I a, I b, bool& ok) {
  if (b.lo <= 0.0 && b.hi >= 0.0) {
    ok = false;
    return a;
  }
  const double x = a.lo / b.lo, y = a.lo / b.hi, z = a.hi / b.lo, w = a.hi / b.hi;
  return outward(np_min(np_min(np_min(x, y), z), w), np_max(np_max(np_max(x, y), z), w), ok);
}

// det of the interval matrix F[r][j] by the left-associated cofactor expansion along row 0
// (compiled_enclosure.py:60-71; feasibility.py:213-216).
__host__ __device__ __forceinline__ I determinant(const I F[3][3], bool& ok) {
  const I a = sub(mul(F[1][1], F[2][2], ok), mul(F[1][2], F[2][1], ok), ok);
  const I b = sub(mul(F[1][2], F[2][0], ok), mul(F[1][0], F[2][2], ok), ok);
  const I c = sub(mul(F[1][0], F[2][1], ok), mul(F[1][1], F[2][0], ok), ok);
  return add(add(mul(F[0][0], a, ok), mul(F[0][1], b, ok), ok), mul(F[0][2], c, ok), ok);
}

// Enclosure of J = det(I + grad u) at the reference point t in [0,1]^3 of a cell with vertex values v
// (stationarity.py:39-56): per derivative direction j, the sum over the four edges along j of
// ((v_b - v_a) / h_j) * w, with the bilinear weight w = prod over the other axes k of (t_k or 1 - t_k) built
// from the point one factor at a time; then F = that + identity and the row-0 cofactor determinant
// F00 (F11 F22 - F12 F21) - F01 (F10 F22 - F12 F20) + F02 (F10 F21 - F11 F20). `ok` turns false where the
// reference raises (a nonfinite or inverted step).
__host__ __device__ inline I point_jd_interval(const double v[8][3], const double h[3], const double t[3], bool& ok) {
  I F[3][3];
  for (int j = 0; j < 3; ++j) {
    I d[3] = {point(0.0), point(0.0), point(0.0)};
    for (int a = 0; a < 8; ++a) {
      if (a & (1 << j)) continue;
      const int b = a | (1 << j);
      I w = point(1.0);
      for (int k = 0; k < 3; ++k) {
        if (k == j) continue;
        const I f = (a & (1 << k)) ? point(t[k]) : sub(point(1.0), point(t[k]), ok);
        w = mul(w, f, ok);
      }
      for (int r = 0; r < 3; ++r) {
        const I q = div(sub(point(v[b][r]), point(v[a][r]), ok), point(h[j]), ok);
        d[r] = add(d[r], mul(q, w, ok), ok);
      }
    }
    for (int r = 0; r < 3; ++r) F[r][j] = add(d[r], point(r == j ? 1.0 : 0.0), ok);
  }
  const I a = sub(mul(F[1][1], F[2][2], ok), mul(F[1][2], F[2][1], ok), ok);
  const I b = sub(mul(F[1][0], F[2][2], ok), mul(F[1][2], F[2][0], ok), ok);
  const I c = sub(mul(F[1][0], F[2][1], ok), mul(F[1][1], F[2][0], ok), ok);
  return add(sub(mul(F[0][0], a, ok), mul(F[0][1], b, ok), ok), mul(F[0][2], c, ok), ok);
}

// The same operations on finite operands (FAST: next_*_finite and plain min/max). Once any step of a cell is
// nonfinite the cell is flagged by both versions; for an unflagged cell every intermediate is finite, where the
// two versions agree bit for bit. The GPU screen uses FAST (fewer registers and branches).
template <bool FAST>
struct Ops {
  static __host__ __device__ __forceinline__ double mn(double a, double b) { return FAST ? (a < b ? a : b) : np_min(a, b); }
  static __host__ __device__ __forceinline__ double mx(double a, double b) { return FAST ? (a > b ? a : b) : np_max(a, b); }
  static __host__ __device__ __forceinline__ I out(double lo, double hi, bool& ok) {
    const I r = FAST ? I{next_down_finite(lo), next_up_finite(hi)} : I{next_down(lo), next_up(hi)};
    ok = ok && is_finite(r.lo) && is_finite(r.hi) && r.lo <= r.hi;
    return r;
  }
  static __host__ __device__ __forceinline__ I add(I a, I b, bool& ok) { return out(a.lo + b.lo, a.hi + b.hi, ok); }
  static __host__ __device__ __forceinline__ I sub(I a, I b, bool& ok) { return out(a.lo - b.hi, a.hi - b.lo, ok); }
  static __host__ __device__ __forceinline__ I mul(I a, I b, bool& ok) {
    const double x = a.lo * b.lo, y = a.lo * b.hi, z = a.hi * b.lo, w = a.hi * b.hi;
    return out(mn(mn(mn(x, y), z), w), mx(mx(mx(x, y), z), w), ok);
  }
};

// Edge-hull margin of one cell (compiled_enclosure.py:193-227): F[r][j] = hull of the four raw edge
// differences along j (widened once, then divided by h_j with one more outward step) plus the identity,
// margin = lo(det F - delta) with the left-associated cofactor determinant. `ok` is false when any step was
// nonfinite or inverted (the cell's flag). Vertex a = x + 2y + 4z; uv[a][r] is component r at vertex a.
template <bool FAST = false>
__host__ __device__ inline double edge_hull_margin(const double uv[8][3], const double h[3], double delta,
                                                   bool& ok) {
  using O = Ops<FAST>;
  const double inv[3] = {pow2_reciprocal(h[0]), pow2_reciprocal(h[1]), pow2_reciprocal(h[2])};
  I F[3][3];
  for (int j = 0; j < 3; ++j) {
    for (int r = 0; r < 3; ++r) {
      double lower = __builtin_inf(), upper = -__builtin_inf();
      bool good = true;
      for (int a = 0; a < 8; ++a) {
        if (a & (1 << j)) continue;
        const int b = a | (1 << j);
        const double d = uv[b][r] - uv[a][r];
        if (!is_finite(d)) good = false;
        lower = O::mn(lower, d);
        upper = O::mx(upper, d);
      }
      I e = O::out(lower, upper, good);
      e = O::out(div_by(e.lo, h[j], inv[j]), div_by(e.hi, h[j], inv[j]), good);
      const double id = r == j ? 1.0 : 0.0;
      F[r][j] = O::add(e, point(id), good);
      ok = ok && good;
    }
  }
  const I a = O::sub(O::mul(F[1][1], F[2][2], ok), O::mul(F[1][2], F[2][1], ok), ok);
  const I b = O::sub(O::mul(F[1][2], F[2][0], ok), O::mul(F[1][0], F[2][2], ok), ok);
  const I c = O::sub(O::mul(F[1][0], F[2][1], ok), O::mul(F[1][1], F[2][0], ok), ok);
  const I J = O::add(O::add(O::mul(F[0][0], a, ok), O::mul(F[0][1], b, ok), ok), O::mul(F[0][2], c, ok), ok);
  const I m = O::out(J.lo - delta, J.hi - delta, ok);
  return m.lo;
}

}  // namespace iv
}  // namespace seqvp
