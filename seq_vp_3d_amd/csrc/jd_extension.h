// T69 face/interior estimate of a cell's Jacobian minimum, for the JD-affine metric
// (jd_curl/three_d/solvers/jd_affine.py:49-188). Host and device compile the same code.
//
// J on a Q1 cell is the triquadratic through its 27 closed-point values J[9 iz + 3 iy + ix] (reference
// nodes 0, 1/2, 1 per axis). The estimate mu is the least of
//   - J on the 9^3 grid {0, 1/8, ..., 1}^3, and
//   - bounded Newton on [0, 1]^3 from the 4 lowest grid points (ties: lower z-major grid index first),
//     30 iterations of at most 40 step halvings each.
// Every iteration fixes a coordinate that sits on a face with the gradient pointing out of the cell,
// diagonalises the Hessian restricted to the free coordinates (identity on the fixed ones, cyclic Jacobi),
// takes the Newton step when every eigenvalue exceeds 1e-12 max(1, |w|max) and the gradient step divided by
// that scale otherwise, and halves it until the clipped point strictly lowers J; a start that cannot lower J
// stops. mu is a value J attains, so it is an estimate for the metric, never a bound.
//
// The grid values, monomial coefficients and polynomial evaluations are separable contractions; they round
// differently from the reference's BLAS product and einsum, so mu agrees with the reference to rounding where
// the start selection and the Newton decisions agree (the reference documents the same latitude for its BLAS
// product, jd_affine.py:50-53).
#pragma once
#include <hip/hip_runtime.h>

namespace seqvp {
namespace t69 {

// The three Lagrange polynomials on 0, 1/2, 1 at t; exact for t = k/8.
__host__ __device__ inline void lagrange3(double t, double L[3]) {
  L[0] = 2.0 * (t - 0.5) * (t - 1.0);
  L[1] = -4.0 * t * (t - 1.0);
  L[2] = 2.0 * t * (t - 0.5);
}

// Least Bernstein coefficient of the cell's triquadratic (bernstein_barrier.py:26-38): per axis the middle
// entry becomes 2 b - (a + c)/2, x first, then y, then z. 2 b and the halving are exact, so the result does
// not depend on floating-point contraction. v holds the 27 closed values and is overwritten with the
// coefficients.
__host__ __device__ inline double bernstein_min_inplace(double v[27]) {
#pragma unroll
  for (int z = 0; z < 3; ++z)
    for (int y = 0; y < 3; ++y) {
      double* r = v + 9 * z + 3 * y;
      r[1] = 2.0 * r[1] - 0.5 * (r[0] + r[2]);
    }
#pragma unroll
  for (int z = 0; z < 3; ++z)
    for (int x = 0; x < 3; ++x) {
      double* r = v + 9 * z + x;
      r[3] = 2.0 * r[3] - 0.5 * (r[0] + r[6]);
    }
#pragma unroll
  for (int y = 0; y < 3; ++y)
    for (int x = 0; x < 3; ++x) {
      double* r = v + 3 * y + x;
      r[9] = 2.0 * r[9] - 0.5 * (r[0] + r[18]);
    }
  double m = v[0];
  for (int p = 1; p < 27; ++p) m = v[p] < m ? v[p] : m;
  return m;
}

// Lagrange values f0, f1, f2 of one axis -> coefficients of 1, t, t^2.
__host__ __device__ inline void to_monomial(double& f0, double& f1, double& f2) {
  const double c1 = -3.0 * f0 + 4.0 * f1 - f2;
  const double c2 = 2.0 * f0 - 4.0 * f1 + 2.0 * f2;
  f1 = c1;
  f2 = c2;
}

// C[9 k + 3 j + i] = coefficient of z^k y^j x^i.
__host__ __device__ inline void monomial(const double J[27], double C[27]) {
  for (int p = 0; p < 27; ++p) C[p] = J[p];
  for (int q = 0; q < 9; ++q) to_monomial(C[3 * q], C[3 * q + 1], C[3 * q + 2]);
  for (int z = 0; z < 3; ++z)
    for (int x = 0; x < 3; ++x) to_monomial(C[9 * z + x], C[9 * z + 3 + x], C[9 * z + 6 + x]);
  for (int q = 0; q < 9; ++q) to_monomial(C[q], C[9 + q], C[18 + q]);
}

__host__ __device__ inline double value(const double C[27], const double x[3]) {
  double s[3];
  for (int k = 0; k < 3; ++k) {
    double r[3];
    for (int j = 0; j < 3; ++j) {
      const double* c = C + 9 * k + 3 * j;
      r[j] = c[0] + x[0] * (c[1] + x[0] * c[2]);
    }
    s[k] = r[0] + x[1] * (r[1] + x[1] * r[2]);
  }
  return s[0] + x[2] * (s[1] + x[2] * s[2]);
}

// Gradient g (x, y, z) and Hessian H (symmetric, full) at x.
__host__ __device__ inline void derivatives(const double C[27], const double x[3], double g[3], double H[3][3]) {
  const double py[3] = {1.0, x[1], x[1] * x[1]}, dy[3] = {0.0, 1.0, 2.0 * x[1]};
  const double pz[3] = {1.0, x[2], x[2] * x[2]}, dz[3] = {0.0, 1.0, 2.0 * x[2]};
  double PP[3], PD[3], DP[3], HP[3], PH[3], DD_[3];
  for (int k = 0; k < 3; ++k) {
    double p[3], d[3], h[3];
    for (int j = 0; j < 3; ++j) {
      const double* c = C + 9 * k + 3 * j;
      p[j] = c[0] + x[0] * (c[1] + x[0] * c[2]);
      d[j] = c[1] + 2.0 * x[0] * c[2];
      h[j] = 2.0 * c[2];
    }
    PP[k] = py[0] * p[0] + py[1] * p[1] + py[2] * p[2];
    PD[k] = py[0] * d[0] + py[1] * d[1] + py[2] * d[2];
    HP[k] = py[0] * h[0] + py[1] * h[1] + py[2] * h[2];
    DP[k] = dy[1] * p[1] + dy[2] * p[2];
    DD_[k] = dy[1] * d[1] + dy[2] * d[2];
    PH[k] = 2.0 * p[2];
  }
  g[0] = pz[0] * PD[0] + pz[1] * PD[1] + pz[2] * PD[2];
  g[1] = pz[0] * DP[0] + pz[1] * DP[1] + pz[2] * DP[2];
  g[2] = dz[1] * PP[1] + dz[2] * PP[2];
  H[0][0] = pz[0] * HP[0] + pz[1] * HP[1] + pz[2] * HP[2];
  H[1][1] = pz[0] * PH[0] + pz[1] * PH[1] + pz[2] * PH[2];
  H[2][2] = 2.0 * PP[2];
  H[0][1] = H[1][0] = pz[0] * DD_[0] + pz[1] * DD_[1] + pz[2] * DD_[2];
  H[0][2] = H[2][0] = dz[1] * PD[1] + dz[2] * PD[2];
  H[1][2] = H[2][1] = dz[1] * DP[1] + dz[2] * DP[2];
}

// Eigen-decomposition of a symmetric 3x3 matrix by the cyclic Jacobi method (Golub and Van Loan, Matrix
// Computations, 4th ed., section 8.5.3): A becomes diagonal (the eigenvalues), V collects the rotations
// (columns are eigenvectors). A rotation is skipped once its off-diagonal entry no longer changes the two
// diagonal entries it couples (Rutishauser's test); at most 50 sweeps.
__host__ __device__ inline void jacobi3(double A[3][3], double V[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) V[i][j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 50; ++sweep) {
    if (A[0][1] == 0.0 && A[0][2] == 0.0 && A[1][2] == 0.0) break;
    for (int pq = 0; pq < 3; ++pq) {
      const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
      const double apq = A[p][q];
      if (apq == 0.0) continue;
      const double app = A[p][p], aqq = A[q][q];
      const double g = 100.0 * fabs(apq);
      if (sweep > 3 && fabs(app) + g == fabs(app) && fabs(aqq) + g == fabs(aqq)) {
        A[p][q] = A[q][p] = 0.0;
        continue;
      }
      const double theta = (aqq - app) / (2.0 * apq);
      double t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
      if (fabs(theta) > 1e150) t = 0.5 / fabs(theta);
      if (theta < 0.0) t = -t;
      const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
      A[p][p] = app - t * apq;
      A[q][q] = aqq + t * apq;
      A[p][q] = A[q][p] = 0.0;
      const int r = 3 - p - q;
      const double arp = A[r][p], arq = A[r][q];
      A[r][p] = A[p][r] = c * arp - s * arq;
      A[r][q] = A[q][r] = s * arp + c * arq;
      for (int i = 0; i < 3; ++i) {
        const double vip = V[i][p], viq = V[i][q];
        V[i][p] = c * vip - s * viq;
        V[i][q] = s * vip + c * viq;
      }
    }
  }
}

// np.clip(v, 0, 1): NaN stays NaN.
__host__ __device__ inline double clip01(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }

// Lowest value bounded Newton reaches from x (jd_affine.py:95-127).
__host__ __device__ inline double bounded_newton(const double C[27], double x[3], int iterations = 30,
                                                 int halvings = 40) {
  double best = value(C, x);
  for (int it = 0; it < iterations; ++it) {
    double g[3], H[3][3];
    derivatives(C, x, g, H);
    bool fixed[3];
    double gf[3];
    for (int i = 0; i < 3; ++i) {
      fixed[i] = (x[i] <= 0.0 && g[i] > 0.0) || (x[i] >= 1.0 && g[i] < 0.0);
      gf[i] = fixed[i] ? 0.0 : g[i];
    }
    double A[3][3], V[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) A[i][j] = (fixed[i] || fixed[j]) ? (i == j ? 1.0 : 0.0) : H[i][j];
    jacobi3(A, V);
    const double w[3] = {A[0][0], A[1][1], A[2][2]};
    double scale = 1.0;
    for (int i = 0; i < 3; ++i) scale = fabs(w[i]) > scale ? fabs(w[i]) : scale;
    const bool definite = w[0] > 1e-12 * scale && w[1] > 1e-12 * scale && w[2] > 1e-12 * scale;
    double d[3];
    if (definite) {
      double t[3];
      for (int j = 0; j < 3; ++j) t[j] = (V[0][j] * gf[0] + V[1][j] * gf[1] + V[2][j] * gf[2]) / w[j];
      for (int i = 0; i < 3; ++i) d[i] = -(V[i][0] * t[0] + V[i][1] * t[1] + V[i][2] * t[2]);
    } else {
      for (int i = 0; i < 3; ++i) d[i] = -gf[i] / scale;
    }
    double step = 1.0;
    bool moved = false;
    for (int h = 0; h < halvings; ++h) {
      const double xt[3] = {clip01(x[0] + step * d[0]), clip01(x[1] + step * d[1]), clip01(x[2] + step * d[2])};
      const double ft = value(C, xt);
      if (ft < best) {
        x[0] = xt[0];
        x[1] = xt[1];
        x[2] = xt[2];
        best = ft;
        moved = true;
        break;
      }
      step *= 0.5;
    }
    if (!moved) break;
  }
  return best;
}

// mu of one cell (jd_affine.py:130-151).
__host__ __device__ inline double face_interior_min(const double J[27]) {
  constexpr int STARTS = 4;
  double bv[STARTS];
  int bi[STARTS];
  for (int s = 0; s < STARTS; ++s) {
    bv[s] = __builtin_inf();
    bi[s] = 0;
  }
  for (int iz = 0; iz < 9; ++iz) {
    double Lz[3];
    lagrange3(0.125 * iz, Lz);
    double P[9];
    for (int q = 0; q < 9; ++q) P[q] = Lz[0] * J[q] + Lz[1] * J[9 + q] + Lz[2] * J[18 + q];
    for (int iy = 0; iy < 9; ++iy) {
      double Ly[3];
      lagrange3(0.125 * iy, Ly);
      double Q[3];
      for (int a = 0; a < 3; ++a) Q[a] = Ly[0] * P[a] + Ly[1] * P[3 + a] + Ly[2] * P[6 + a];
      for (int ix = 0; ix < 9; ++ix) {
        double Lx[3];
        lagrange3(0.125 * ix, Lx);
        const double v = Lx[0] * Q[0] + Lx[1] * Q[1] + Lx[2] * Q[2];
        const int gi = 81 * iz + 9 * iy + ix;
        // sorted insertion; an equal value goes after the ones already held (a stable order)
        if (v < bv[3]) {
          bv[3] = v;
          bi[3] = gi;
#pragma unroll
          for (int s = 3; s > 0; --s) {
            if (bv[s] < bv[s - 1]) {
              const double tv = bv[s];
              bv[s] = bv[s - 1];
              bv[s - 1] = tv;
              const int ti = bi[s];
              bi[s] = bi[s - 1];
              bi[s - 1] = ti;
            }
          }
        }
      }
    }
  }
  double C[27];
  monomial(J, C);
  double mu = bv[0];
  for (int s = 0; s < STARTS; ++s) {
    double x[3] = {0.125 * (bi[s] % 9), 0.125 * ((bi[s] / 9) % 9), 0.125 * (bi[s] / 81)};
    const double f = bounded_newton(C, x);
    mu = f < mu ? f : mu;
  }
  return mu;
}

}  // namespace t69
}  // namespace seqvp
