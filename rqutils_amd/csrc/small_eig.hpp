// Small dense eigenproblems of the LOBPCG, evaluated by one device thread in
// FP64: the reference's closed forms for 2x2 and 3x3 Hermitian matrices
// (rqutils/ground_locg.py:424-471) and the generalized 3x3 Rayleigh-Ritz step
// of the fused iteration (Cholesky of the Gram matrix + complex Jacobi).
#pragma once

#include "common.hpp"

namespace sqdk {

// ---- small closed forms, as ground_locg.py computes them -------------------

template <int N>
__device__ inline void jacobi_lowest(double2 a[3][3], double& lam, double2 vec[3]);

// Lowest eigenpair of a 2x2 Hermitian matrix (ground_locg.py:424-435). Where
// the closed form divides by zero (d0 - lam + m10 == 0, e.g. a diagonal
// matrix whose lower entry is the smaller) the reference returns NaN; here the
// pair then comes from the Jacobi rotation instead.
__device__ inline void eig2_reference(const double2 m[2][2], double& lam, double2 vec[2]) {
  const double d0 = m[0][0].x, d1 = m[1][1].x;
  const double det = d0 * d1 - cabs2(m[1][0]);
  const double tr = d0 + d1;
  lam = (tr - sqrt(fmax(tr * tr - 4.0 * det, 0.0))) * 0.5;
  const double2 num = make_double2(d1 - lam + m[1][0].x, -m[1][0].y);
  const double2 den = make_double2(d0 - lam + m[1][0].x, m[1][0].y);
  const double dd = cabs2(den);
  const double2 e0 = make_double2((num.x * den.x + num.y * den.y) / dd,
                                  (num.y * den.x - num.x * den.y) / dd);
  const double nrm = sqrt(e0.x * e0.x + e0.y * e0.y + 1.0);
  if (dd > 0.0 && isfinite(nrm)) {
    vec[0] = cscale(e0, 1.0 / nrm);
    vec[1] = make_double2(-1.0 / nrm, 0.0);
    return;
  }
  double2 a[3][3] = {{m[0][0], m[0][1], make_double2(0.0, 0.0)},
                     {m[1][0], m[1][1], make_double2(0.0, 0.0)},
                     {make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0)}};
  double2 v3[3];
  jacobi_lowest<2>(a, lam, v3);
  vec[0] = v3[0];
  vec[1] = v3[1];
}

__device__ inline void cross3(const double2 a[3], const double2 b[3], double2 e[3]) {
  e[0] = csub(cmul(a[1], b[2]), cmul(a[2], b[1]));
  e[1] = csub(cmul(a[2], b[0]), cmul(a[0], b[2]));
  e[2] = csub(cmul(a[0], b[1]), cmul(a[1], b[0]));
}

// Lowest eigenpair of a 3x3 Hermitian matrix by Cardano's method
// (J. Kopp, Int. J. Mod. Phys. C 19, 523 (2008); ground_locg.py:438-471):
// characteristic coefficients, the trigonometric root, and the eigenvector as
// the conjugated cross product of columns 1 and 2 of (A - lambda I).
//
// Guards where the reference's formulas give NaN or a wrong vector: rounding
// can make the two square-root arguments slightly negative (clamped at 0), and
// columns 1 and 2 of A - lambda I can be parallel, zero or pure rounding noise
// (a diagonal matrix, a double eigenvalue), so their cross product vanishes or
// points anywhere. The eigenvector is then the conjugated cross product of the
// strongest column pair (each row of the singular Hermitian matrix
// A - lambda I is orthogonal to it), and when A - lambda I has rank one or
// zero, the lowest Jacobi eigenvector.
__device__ inline void eig3_reference(const double2 m[3][3], double& lam, double2 vec[3]) {
  const double d0 = m[0][0].x, d1 = m[1][1].x, d2 = m[2][2].x;
  const double o10 = cabs2(m[1][0]), o20 = cabs2(m[2][0]), o21 = cabs2(m[2][1]);
  const double c2 = -(d0 + d1 + d2);
  const double c1 = (d0 * d2 + d1 * d0 + d2 * d1) - (o10 + o20 + o21);
  double c0 = d0 * o21 + d1 * o20 + d2 * o10;
  c0 -= d0 * d1 * d2;
  c0 -= 2.0 * cmul(cmul(m[0][2], m[1][0]), m[2][1]).x;
  const double p = fmax(c2 * c2 - 3.0 * c1, 0.0);
  const double q = -13.5 * c0 - c2 * c2 * c2 + 4.5 * c2 * c1;
  const double phi =
      atan2(sqrt(fmax(27.0 * (0.25 * c1 * c1 * (p - c1) + c0 * (q + 6.75 * c0)), 0.0)), q) / 3.0;
  const double cphi = cos(phi), sphi = sin(phi);
  const double s3 = sqrt(3.0);
  const double xmin = fmin(fmin(2.0 * cphi, -cphi - s3 * sphi), -cphi + s3 * sphi);
  lam = sqrt(p) / 3.0 * xmin - c2 / 3.0;
  // columns of A - lambda I
  double2 col[3][3];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i) col[j][i] = m[i][j];
#pragma unroll
  for (int j = 0; j < 3; ++j) col[j][j].x -= lam;
  double fro = 0.0;
#pragma unroll
  for (int j = 0; j < 3; ++j) fro += cabs2(col[j][0]) + cabs2(col[j][1]) + cabs2(col[j][2]);
  double2 e[3], e01[3], e02[3];
  cross3(col[1], col[2], e);
  cross3(col[0], col[1], e01);
  cross3(col[0], col[2], e02);
  double en = cabs2(e[0]) + cabs2(e[1]) + cabs2(e[2]);
  const double n01 = cabs2(e01[0]) + cabs2(e01[1]) + cabs2(e01[2]);
  const double n02 = cabs2(e02[0]) + cabs2(e02[1]) + cabs2(e02[2]);
  const double best = fmax(en, fmax(n01, n02));
  // |cross| below 1e-10 |A - lambda I|_F^2 for every pair: rank one or zero
  if (!(isfinite(best) && isfinite(fro) && best > 1e-20 * fro * fro)) {
    double2 a[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) a[i][j] = m[i][j];
    jacobi_lowest<3>(a, lam, vec);
    return;
  }
  // The reference's pair unless it is far weaker than the best one (then it
  // is mostly rounding noise: for diag(3, -2, 1) column 1 is 4e-16 and the
  // reference's vector points along e0 instead of e1).
  if (!(en >= 1e-6 * best)) {
    const bool use01 = n01 >= n02;
#pragma unroll
    for (int i = 0; i < 3; ++i) e[i] = use01 ? e01[i] : e02[i];
    en = use01 ? n01 : n02;
  }
  const double nrm = sqrt(en);
  for (int i = 0; i < 3; ++i) vec[i] = cscale(cconj(e[i]), 1.0 / nrm);
}

// ---- robust small eigensolver for the fused Rayleigh-Ritz ------------------
//
// The sizes are template parameters and every loop is unrolled, so the 3x3
// arrays live in registers (indexed by a runtime size they would go to
// scratch memory, which made one Rayleigh-Ritz step cost tens of microseconds).

// Cyclic complex Jacobi on an N x N (N <= 3) Hermitian matrix: each rotation
// first removes the phase of a_pq, then applies the real Jacobi rotation.
// Returns the lowest eigenvalue and its unit eigenvector.
template <int N>
__device__ inline void jacobi_lowest(double2 a[3][3], double& lam, double2 vec[3]) {
  double2 v[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) v[i][j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = 0.0, diag = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      diag += a[i][i].x * a[i][i].x;
#pragma unroll
      for (int j = i + 1; j < N; ++j) off += cabs2(a[i][j]);
    }
    if (off <= 1e-32 * diag || off == 0.0) break;
#pragma unroll
    for (int p = 0; p < N; ++p) {
#pragma unroll
      for (int q = p + 1; q < N; ++q) {
        const double b = sqrt(cabs2(a[p][q]));
        if (b == 0.0) continue;
        // phase: D = diag(.., e^{-i phi} at q) with a_pq = b e^{i phi}
        const double2 ph = make_double2(a[p][q].x / b, -a[p][q].y / b);  // e^{-i phi}
#pragma unroll
        for (int k = 0; k < N; ++k) {
          a[k][q] = cmul(a[k][q], ph);          // column q times e^{-i phi}
          v[k][q] = cmul(v[k][q], ph);
        }
#pragma unroll
        for (int k = 0; k < N; ++k) a[q][k] = cmul(a[q][k], cconj(ph));  // row q times e^{i phi}
        // now a_pq = b is real; real rotation zeroing it
        const double app = a[p][p].x, aqq = a[q][q].x;
        const double tau = (aqq - app) / (2.0 * b);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const double2 akp = a[k][p], akq = a[k][q];
          a[k][p] = csub(cscale(akp, c), cscale(akq, s));
          a[k][q] = cadd(cscale(akp, s), cscale(akq, c));
          const double2 vkp = v[k][p], vkq = v[k][q];
          v[k][p] = csub(cscale(vkp, c), cscale(vkq, s));
          v[k][q] = cadd(cscale(vkp, s), cscale(vkq, c));
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const double2 apk = a[p][k], aqk = a[q][k];
          a[p][k] = csub(cscale(apk, c), cscale(aqk, s));
          a[q][k] = cadd(cscale(apk, s), cscale(aqk, c));
        }
        a[p][q] = make_double2(0.0, 0.0);
        a[q][p] = make_double2(0.0, 0.0);
      }
    }
  }
  // the lowest diagonal entry (the first of equal ones) and its column
  lam = a[0][0].x;
#pragma unroll
  for (int i = 0; i < 3; ++i) vec[i] = i < N ? v[i][0] : make_double2(0.0, 0.0);
#pragma unroll
  for (int j = 1; j < N; ++j) {
    if (a[j][j].x < lam) {
      lam = a[j][j].x;
#pragma unroll
      for (int i = 0; i < N; ++i) vec[i] = v[i][j];
    }
  }
}

// Lower Cholesky factor of an N x N Hermitian matrix; false if a pivot^2 is
// not above `floor` (the matrix has unit diagonal, so pivots are sines).
template <int N>
__device__ inline bool cholesky(const double2 g[3][3], double2 l[3][3], double floor) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) l[i][j] = make_double2(0.0, 0.0);
#pragma unroll
  for (int j = 0; j < N; ++j) {
    double s = g[j][j].x;
#pragma unroll
    for (int k = 0; k < j; ++k) s -= cabs2(l[j][k]);
    if (!(s > floor)) return false;
    const double ljj = sqrt(s);
    l[j][j] = make_double2(ljj, 0.0);
#pragma unroll
    for (int i = j + 1; i < N; ++i) {
      double2 t = g[i][j];
#pragma unroll
      for (int k = 0; k < j; ++k) t = csub(t, cmul(l[i][k], cconj(l[j][k])));
      l[i][j] = cscale(t, 1.0 / ljj);
    }
  }
  return true;
}

// The generalized eigenproblem M c = lam G c on the first N basis vectors,
// after the diagonal scaling dsc: Cholesky G = L L^H (false if G is not
// numerically positive definite), C = L^-1 M L^-H, its lowest eigenpair by
// Jacobi, c = L^-H y with the scaling undone.
template <int N>
__device__ inline bool rr_solve(const double2 G[3][3], const double2 M[3][3], const double dsc[3],
                                double floor, double2 c[3], double& lam) {
  double2 Gs[3][3], Ms[3][3], L[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const bool in = i < N && j < N;
      Gs[i][j] = in ? cscale(G[i][j], dsc[i] * dsc[j]) : make_double2(0.0, 0.0);
      Ms[i][j] = in ? cscale(M[i][j], dsc[i] * dsc[j]) : make_double2(0.0, 0.0);
    }
  if (!cholesky<N>(Gs, L, floor)) return false;
  // Linv = L^{-1} (lower triangular)
  double2 Li[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Li[i][j] = make_double2(0.0, 0.0);
#pragma unroll
  for (int j = 0; j < N; ++j) {
    Li[j][j] = make_double2(1.0 / L[j][j].x, 0.0);
#pragma unroll
    for (int i = j + 1; i < N; ++i) {
      double2 t = make_double2(0.0, 0.0);
#pragma unroll
      for (int k = j; k < i; ++k) t = cadd(t, cmul(L[i][k], Li[k][j]));
      Li[i][j] = cscale(make_double2(-t.x, -t.y), 1.0 / L[i][i].x);
    }
  }
  // Cm = Linv Ms Linv^H
  double2 T[3][3], Cm[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double2 t = make_double2(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < N; ++k) t = cadd(t, cmul(Li[i][k], Ms[k][j]));
      T[i][j] = t;
    }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double2 t = make_double2(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < N; ++k) t = cadd(t, cmul(T[i][k], cconj(Li[j][k])));
      Cm[i][j] = t;
    }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    Cm[i][i].y = 0.0;
#pragma unroll
    for (int j = 0; j < i; ++j) {
      const double2 h = cscale(cadd(Cm[i][j], cconj(Cm[j][i])), 0.5);
      Cm[i][j] = h;
      Cm[j][i] = cconj(h);
    }
  }
  double2 y[3];
  jacobi_lowest<N>(Cm, lam, y);
  // c = Linv^H y, then undo the diagonal scaling
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double2 t = make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < N; ++k) t = cadd(t, cmul(cconj(Li[k][i]), y[k]));
    c[i] = i < N ? cscale(t, dsc[i]) : make_double2(0.0, 0.0);
  }
  return true;
}

// Generalized Rayleigh-Ritz over S = (X, R, P) from the 13 fused Gram sums
// g[0..5] = XX RR PP XR XP RP, g[6..11] = X.AX R.AR P.AP X.AR X.AP R.AP:
// minimise c^H M c / c^H G c. Returns 0 and the update coefficients
// U = (c0, c1, c2, c1 s, c2 s, lam) with s normalising P' = c1 R + c2 P, or 2
// when R is numerically parallel to X. P is used only if pvalid; if the
// 3-vector Gram matrix is not numerically positive definite, (X, R) is used.
__device__ inline int fused_rr(const double2* g, bool pvalid, int precision, double2 U[6],
                               bool& pvalid_out) {
  // G_ij = (S_i, S_j), M_ij = (S_i, A S_j) over S = (X, R, P)
  double2 G[3][3], M[3][3];
  G[0][0] = g[0]; G[1][1] = g[1]; G[2][2] = g[2];
  G[0][1] = g[3]; G[0][2] = g[4]; G[1][2] = g[5];
  M[0][0] = g[6]; M[1][1] = g[7]; M[2][2] = g[8];
  M[0][1] = g[9]; M[0][2] = g[10]; M[1][2] = g[11];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    G[i][i].y = 0.0;
    M[i][i].y = 0.0;
#pragma unroll
    for (int j = 0; j < i; ++j) {
      G[i][j] = cconj(G[j][i]);
      M[i][j] = cconj(M[j][i]);
    }
  }
  const bool three = pvalid && G[2][2].x > 0.0;
  const double dsc[3] = {1.0 / sqrt(G[0][0].x), 1.0 / sqrt(G[1][1].x),
                         three ? 1.0 / sqrt(G[2][2].x) : 0.0};
  const double floor = precision == SQDK_C128 ? 1e-14 : 1e-8;
  double2 c[3];
  double lam;
  bool ok = three && rr_solve<3>(G, M, dsc, floor, c, lam);
  if (!ok) ok = rr_solve<2>(G, M, dsc, floor, c, lam);
  if (!ok) return 2;  // R is numerically parallel to X: nothing left to improve
  // fix the phase so the X coefficient is real and non-negative
  const double a0 = sqrt(cabs2(c[0]));
  if (a0 > 0.0) {
    const double2 ph = make_double2(c[0].x / a0, -c[0].y / a0);
#pragma unroll
    for (int i = 0; i < 3; ++i) c[i] = cmul(c[i], ph);
  }
  // P' = c1 R + c2 P, normalised with the Gram matrix
  double pn2 = 0.0;
#pragma unroll
  for (int i = 1; i < 3; ++i)
#pragma unroll
    for (int j = 1; j < 3; ++j) pn2 += cmul(cconj(c[i]), cmul(G[i][j], c[j])).x;
  const double pscale = pn2 > 0.0 ? 1.0 / sqrt(pn2) : 0.0;
  U[0] = c[0];
  U[1] = c[1];
  U[2] = c[2];
  U[3] = cscale(c[1], pscale);
  U[4] = cscale(c[2], pscale);
  U[5] = make_double2(lam, 0.0);
  pvalid_out = pscale > 0.0;
  return 0;
}

}  // namespace sqdk
