// Host-side IWP system matrices (shared by the C ABI and the CPU build of
// the kernels).
#pragma once

#include <cmath>
#include <vector>

namespace pdq {

// ------------------------------------------------- IWP system matrices
// utilities.py:57-71: A = flip(Pascal) (binomials C(nu-i, nu-j)), Q = lower
// Cholesky factor of the flipped Hilbert matrix 1/(2nu+1-i-j) with positive
// diagonal.  Q is formed as in the reference: Kahan's recurrence for the
// Hilbert Cholesky factor (cholesky_util.py:106-176), row flip, one
// Householder triangularisation and a sign fix.
inline double binom(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (double)(n - k + i) / (double)i;
  return r;
}

constexpr int kIwpMaxN = 8;  // nu + 1 <= 8

// Into caller arrays of (nu+1)^2 doubles; nu <= 7.  No heap allocation.
inline void iwp_matrices_into(int nu, double* A, double* Q) {
  const int n = nu + 1;
  for (int i = 0; i < n * n; ++i) A[i] = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) A[i * n + j] = binom(nu - i, nu - j);
  // Kahan: U upper with H = U^T U ... columns scaled by sqrt(2i+1) f_j
  double f[kIwpMaxN], U[kIwpMaxN * kIwpMaxN], Lh[kIwpMaxN * kIwpMaxN], M[kIwpMaxN * kIwpMaxN];
  f[0] = 1.0;
  for (int k = 1; k < n; ++k) f[k] = (((f[k - 1] / k) * (2.0 * k)) / k) * (2.0 * k + 1.0);
  for (int i = 0; i < n * n; ++i) U[i] = Lh[i] = 0.0;
  for (int i = 0; i < n; ++i) U[i * n + i] = 1.0;
  for (int j = 1; j < n; ++j)
    for (int i = j - 1; i >= 0; --i) U[i * n + j] = (U[(i + 1) * n + j] / (double)(j - i)) * (double)(i + j + 2);
  // lower factor = U^T scaled
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const double u = U[i * n + j] * (std::sqrt(2.0 * i + 1.0) * (1.0 / f[j]));
      if (j >= i) Lh[j * n + i] = u;
    }
  // M = (row-flipped Lh)^T, triangularise, transpose back
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) M[i * n + j] = Lh[(n - 1 - j) * n + i];
  for (int i = 0; i < n - 1; ++i) {
    double ss = 0.0;
    for (int r = i + 1; r < n; ++r) ss += M[r * n + i] * M[r * n + i];
    if (ss == 0.0) continue;
    const double alpha = M[i * n + i];
    const double beta = -std::copysign(std::hypot(alpha, std::sqrt(ss)), alpha);
    const double tau = (beta - alpha) / beta, scal = 1.0 / (alpha - beta);
    for (int r = i + 1; r < n; ++r) M[r * n + i] *= scal;
    for (int j = i + 1; j < n; ++j) {
      double w = M[i * n + j];
      for (int r = i + 1; r < n; ++r) w += M[r * n + i] * M[r * n + j];
      w *= tau;
      M[i * n + j] -= w;
      for (int r = i + 1; r < n; ++r) M[r * n + j] -= M[r * n + i] * w;
    }
    M[i * n + i] = beta;
  }
  for (int i = 0; i < n * n; ++i) Q[i] = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) Q[i * n + j] = M[j * n + i];
  for (int j = 0; j < n; ++j) {
    const double s = Q[j * n + j] < 0 ? -1.0 : 1.0;
    for (int i = 0; i < n; ++i) Q[i * n + j] *= s;
  }
}

inline void iwp_matrices(int nu, std::vector<double>& A, std::vector<double>& Q) {
  const int n = nu + 1;
  A.assign(n * n, 0.0);
  Q.assign(n * n, 0.0);
  iwp_matrices_into(nu, A.data(), Q.data());
}

// Dense layout into caller arrays of ((nu+1) d)^2 doubles (see iwp_dense).
inline void iwp_dense_into(int nu, int d, double base, double* A, double* Q) {
  double A1[kIwpMaxN * kIwpMaxN], Q1[kIwpMaxN * kIwpMaxN];
  iwp_matrices_into(nu, A1, Q1);
  const int N = nu + 1, n = N * d;
  for (size_t i = 0; i < (size_t)n * n; ++i) A[i] = Q[i] = 0.0;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j)
      for (int k = 0; k < d; ++k) {
        A[(size_t)(i * d + k) * n + j * d + k] = A1[i * N + j];
        Q[(size_t)(i * d + k) * n + j * d + k] = Q1[i * N + j] * base;
      }
}

// Dense layout: A = kron(A1, I_d), Q = kron(Q1, base * I_d), derivative-major.
inline void iwp_dense(int nu, int d, double base, std::vector<double>& A, std::vector<double>& Q) {
  const size_t n = (size_t)(nu + 1) * d;
  A.assign(n * n, 0.0);
  Q.assign(n * n, 0.0);
  iwp_dense_into(nu, d, base, A.data(), Q.data());
}

}  // namespace pdq
