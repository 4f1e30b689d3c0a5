/* T75 Bernstein prescreen restated in plain C (oracle; test infrastructure only).
 *
 * Follows jd_curl/three_d/certificate_prescreen.py:60-121 operation for operation: binary64, every
 * +, -, * rounded separately in the stated order (compiled with -ffp-contract=off, no -ffast-math),
 * so its values are the reference kernel's bit for bit. The device kernel must equal these bits.
 */
#include <math.h>
#include <stdint.h>

static const double UNIT = 0x1p-53;
static const double ETA = 1e-290;
static const double SQRT27_UP = 5.196152422706633;

static double gamma_n(double n) { return n * UNIT / (1.0 - n * UNIT); }

void oracle_prescreen(const double* u, const double* D, const int64_t* cells, int64_t ncells, int64_t nx,
                      int64_t ny, int64_t nz, double* out) {
  const int64_t cx = nx - 1, cy = ny - 1;
  int64_t offsets[8];
  for (int a = 0; a < 8; ++a) {
    const int64_t x = a & 1, y = (a >> 1) & 1, z = (a >> 2) & 1; /* VERTICES: a = x + 2y + 4z */
    offsets[a] = (z * ny + y) * nx + x;
  }
  const double g12 = gamma_n(12.0), g6 = gamma_n(6.0), g3 = gamma_n(3.0);
  (void)nz;
  for (int64_t i = 0; i < ncells; ++i) {
    const int64_t c = cells ? cells[i] : i;
    const int64_t x0 = c % cx, y0 = (c / cx) % cy, z0 = c / (cx * cy);
    const int64_t base = (z0 * ny + y0) * nx + x0;
    double b[27], F[3][3];
    double sigma = 0.0, phi = 0.0, big = 0.0;
    int finite = 1;
    for (int p = 0; p < 27; ++p) {
      for (int r = 0; r < 3; ++r) {
        for (int j = 0; j < 3; ++j) {
          double acc = 0.0, mag = 0.0;
          for (int a = 0; a < 8; ++a) {
            const double term = u[3 * (base + offsets[a]) + r] * D[(j * 27 + p) * 8 + a];
            acc += term;
            mag += fabs(term);
          }
          F[r][j] = acc;
          if (mag > sigma) sigma = mag;
        }
        F[r][r] += 1.0;
      }
      for (int r = 0; r < 3; ++r)
        for (int j = 0; j < 3; ++j)
          if (fabs(F[r][j]) > phi) phi = fabs(F[r][j]);
      const double v = (F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) +
                        F[0][1] * (F[1][2] * F[2][0] - F[1][0] * F[2][2])) +
                       F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]);
      b[p] = v;
      if (fabs(v) > big) big = fabs(v);
    }
    const int strides[3] = {1, 3, 9};
    for (int s = 0; s < 3; ++s) {
      const int stride = strides[s];
      for (int p = 0; p < 27; ++p)
        if ((p / stride) % 3 == 1) b[p] = 2.0 * b[p] - 0.5 * (b[p - stride] + b[p + stride]);
    }
    const double e = g12 * (sigma + phi) + ETA;
    const double P = phi + e;
    const double ej = SQRT27_UP * e * (3.0 * P * P + 3.0 * P * e + e * e) + 6.0 * g6 * P * P * P;
    const double bound = 1.01 * (27.0 * ej + 135.0 * g3 * big) + ETA;
    double low = INFINITY;
    for (int p = 0; p < 27; ++p) {
      const double w = b[p];
      if (!(w == w) || w == INFINITY || w == -INFINITY)
        finite = 0;
      else if (w < low)
        low = w;
    }
    if (!(bound == bound) || bound == INFINITY || !(sigma == sigma) || sigma == INFINITY) finite = 0;
    out[i] = finite ? low - bound : NAN;
  }
}
