// Edge-hull margins of four cells at once on the host CPU (AVX2). Every lane performs the operations of
// iv::edge_hull_margin<true> in the same order -- the same IEEE additions, subtractions, products and quotients,
// the same minimum/maximum selections, the same one-ulp outward steps on the bits -- so each lane's margin and
// flag equal the scalar ones bit for bit. Host code only; callers check avx2_available() first.
#pragma once

#include <cstdint>

#include "interval.h"

namespace seqvp {
namespace iv4 {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef long l4 __attribute__((ext_vector_type(4)));   // the lane masks of d4 comparisons

#define SEQVP_AVX2 static inline __attribute__((always_inline, target("avx2")))

SEQVP_AVX2 l4 bits(d4 x) { return __builtin_bit_cast(l4, x); }
SEQVP_AVX2 d4 value(l4 b) { return __builtin_bit_cast(d4, b); }
SEQVP_AVX2 d4 sel(l4 m, d4 a, d4 b) { return value((m & bits(a)) | (~m & bits(b))); }
SEQVP_AVX2 l4 finite(d4 x) { return (x - x) == 0.0; }
// a < b ? a : b and a > b ? a : b, lane by lane (the scalar FAST minimum and maximum)
SEQVP_AVX2 d4 mn(d4 a, d4 b) { return sel(a < b, a, b); }
SEQVP_AVX2 d4 mx(d4 a, d4 b) { return sel(a > b, a, b); }
// next_up_finite: 0 (either sign) steps to the smallest subnormal, otherwise the bits move by one towards +inf
SEQVP_AVX2 d4 up(d4 x) {
  const l4 step = ((x > 0.0) & 2) - 1;
  const l4 zero = x == 0.0;
  return value((zero & (l4)1) | (~zero & (bits(x) + step)));
}
SEQVP_AVX2 d4 down(d4 x) { return -up(-x); }

struct I4 {
  d4 lo, hi;
};

SEQVP_AVX2 I4 out(d4 lo, d4 hi, l4& ok) {
  const I4 r{down(lo), up(hi)};
  ok = ok & finite(r.lo) & finite(r.hi) & (r.lo <= r.hi);
  return r;
}
SEQVP_AVX2 I4 add(I4 a, I4 b, l4& ok) { return out(a.lo + b.lo, a.hi + b.hi, ok); }
SEQVP_AVX2 I4 sub(I4 a, I4 b, l4& ok) { return out(a.lo - b.hi, a.hi - b.lo, ok); }
SEQVP_AVX2 I4 mul(I4 a, I4 b, l4& ok) {
  const d4 x = a.lo * b.lo, y = a.lo * b.hi, z = a.hi * b.lo, w = a.hi * b.hi;
  return out(mn(mn(mn(x, y), z), w), mx(mx(mx(x, y), z), w), ok);
}

// uv[a][r] holds component r at vertex a (a = x + 2y + 4z) of the four cells; returns the four margins, `ok`
// the four flags' complements (all bits set = unflagged).
SEQVP_AVX2 d4 edge_hull_margin(const d4 uv[8][3], const double h[3], double delta, l4& ok) {
  I4 F[3][3];
  for (int j = 0; j < 3; ++j) {
    for (int r = 0; r < 3; ++r) {
      d4 lower = __builtin_inf(), upper = -__builtin_inf();
      l4 good = (l4)-1;
      for (int a = 0; a < 8; ++a) {
        if (a & (1 << j)) continue;
        const int b = a | (1 << j);
        const d4 d = uv[b][r] - uv[a][r];
        good = good & finite(d);
        lower = mn(lower, d);
        upper = mx(upper, d);
      }
      I4 e = out(lower, upper, good);
      e = out(e.lo / h[j], e.hi / h[j], good);
      const d4 id = r == j ? 1.0 : 0.0;
      F[r][j] = add(e, I4{id, id}, good);
      ok = ok & good;
    }
  }
  const I4 a = sub(mul(F[1][1], F[2][2], ok), mul(F[1][2], F[2][1], ok), ok);
  const I4 b = sub(mul(F[1][2], F[2][0], ok), mul(F[1][0], F[2][2], ok), ok);
  const I4 c = sub(mul(F[1][0], F[2][1], ok), mul(F[1][1], F[2][0], ok), ok);
  const I4 J = add(add(mul(F[0][0], a, ok), mul(F[0][1], b, ok), ok), mul(F[0][2], c, ok), ok);
  const I4 m = out(J.lo - delta, J.hi - delta, ok);
  return m.lo;
}

#undef SEQVP_AVX2

inline bool avx2_available() {
  static const bool yes = __builtin_cpu_supports("avx2");
  return yes;
}

}  // namespace iv4
}  // namespace seqvp
