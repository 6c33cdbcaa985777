// ===========================================================================
// probdiffeq CPU FP64 oracle -- TEST INFRASTRUCTURE ONLY.
//
// A line-by-line C++ restatement of the reference's per-step probabilistic
// IVP solver path (JAX code, not runnable here).  Only tests/, the smoke()
// check and bench.py's CPU-baseline leg may load this library, and only as
// the checker.  The product (probdiffeq_amd/, HIP kernels) never links it.
//
// Every function names the reference file:line it restates (paths are
// relative to the reference checkout).
//
// Parity status of each part (what pins it) is listed in tests/ and in the
// design notes:
//   qr_r / sum_of_sqrtm_factors / revert_conditional -- numpy LAPACK geqrf
//       signs, Kalman closed forms (tests/test_oracle_algebra.py)
//   IWP system matrices / transition -- closed forms + mpmath Cholesky
//   logpdf -- scipy.stats.multivariate_normal
//   controller -- PI(1,0) == I KAT
//   whole solves -- scipy DOP853 rtol 1e-13, convergence, equivalences
//       (dense == isotropic, vmap(dense) == blockdiag, fixed-point ==
//       fixed-interval, fixed grid == adaptive grid, truncated save_at)
//   LML -- brute-force joint Gaussian density on small problems
//   Taylor init -- sympy derivatives of the vector fields
//   error estimators (error_state_std, re_linearize_before_error,
//       error_per_unit_step) -- one-step closed forms, DOP853 accuracy
//       (tests/test_oracle_error.py)
//   isotropic ts1 (exact trace) -- dense == isotropic for f = lambda u + t
//       (tests/test_oracle_solves.py)
//   exp_gram / exponential prior -- scipy expm matrix fraction, OU closed
//       form, zero drift == IWP, expm(M dt) block (tests/test_oracle_prior.py)
//   off-grid marginals, terminal-value LML -- Gaussian prediction / RTS
//       closed forms, Gaussian log-density (tests/test_oracle_offgrid.py)
//   posterior samples -- smoothing means and covariances through the
//       sampler's affine dependence on its draws (tests/test_oracle_sample.py)
//   Monte Carlo Jacobian handlers -- the estimator against numpy's v^T J
//       products on the same probes, exact for diagonal Jacobians, the
//       probes' balance, the limit of many probes, and rejected attempts
//       reusing their probes (tests/test_oracle_jacobian.py); the probe
//       stream is this project's generator, not JAX's threefry, so the
//       reference's draws are not reproduced
// ===========================================================================
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

// The scalar type of the whole restatement, including the C API.  The
// oracle proper is built with double (strict and FMA-contracted); the
// "extended" build uses long double (64-bit mantissa) as a higher-precision
// reference against which the FP64 results' own rounding errors are measured.
#ifndef PDQO_REAL
#define PDQO_REAL double
#endif
using Real = PDQO_REAL;

namespace {

// ---------------------------------------------------------------- matrices
struct Mat {
  int r = 0, c = 0;
  std::vector<Real> a;
  Mat() = default;
  Mat(int r_, int c_) : r(r_), c(c_), a((size_t)r_ * c_, 0.0) {}
  Real& operator()(int i, int j) { return a[(size_t)i * c + j]; }
  Real operator()(int i, int j) const { return a[(size_t)i * c + j]; }
  size_t size() const { return a.size(); }
};
using Vec = std::vector<Real>;

Mat eye(int n) {
  Mat m(n, n);
  for (int i = 0; i < n; ++i) m(i, i) = 1.0;
  return m;
}

Mat transpose(const Mat& A) {
  Mat T(A.c, A.r);
  for (int i = 0; i < A.r; ++i)
    for (int j = 0; j < A.c; ++j) T(j, i) = A(i, j);
  return T;
}

// jnp matmul: acc = sum_k A(i,k) B(k,j), k ascending.
Mat matmul(const Mat& A, const Mat& B) {
  Mat C(A.r, B.c);
  for (int i = 0; i < A.r; ++i)
    for (int j = 0; j < B.c; ++j) {
      Real acc = 0.0;
      for (int k = 0; k < A.c; ++k) acc += A(i, k) * B(k, j);
      C(i, j) = acc;
    }
  return C;
}

Mat add(const Mat& A, const Mat& B) {
  Mat C = A;
  for (size_t i = 0; i < C.a.size(); ++i) C.a[i] += B.a[i];
  return C;
}

Mat sub(const Mat& A, const Mat& B) {
  Mat C = A;
  for (size_t i = 0; i < C.a.size(); ++i) C.a[i] -= B.a[i];
  return C;
}

Mat vstack(const Mat& A, const Mat& B) {
  Mat C(A.r + B.r, A.c);
  for (int i = 0; i < A.r; ++i)
    for (int j = 0; j < A.c; ++j) C(i, j) = A(i, j);
  for (int i = 0; i < B.r; ++i)
    for (int j = 0; j < B.c; ++j) C(A.r + i, j) = B(i, j);
  return C;
}

// Row scaling s[:, None] * A.
Mat rowscale(const Vec& s, const Mat& A, bool absval = false) {
  Mat C = A;
  for (int i = 0; i < A.r; ++i) {
    Real f = absval ? std::fabs(s[i]) : s[i];
    for (int j = 0; j < A.c; ++j) C(i, j) = f * A(i, j);
  }
  return C;
}

Mat scale(Real f, const Mat& A) {
  Mat C = A;
  for (auto& x : C.a) x = f * x;
  return C;
}

Real fro(const Mat& A) {
  Real s = 0.0;
  for (Real x : A.a) s += x * x;
  return std::sqrt(s);
}

// backend/linalg.py:8-37 qr_r = jnp.linalg.qr(mode="r"): LAPACK dgeqr2
// Householder convention (dlarfg: beta = -sign(alpha) * ||(alpha, x)||,
// tau = 0 and no reflection when the sub-column is zero).
Mat qr_r(const Mat& M) {
  Mat A = M;
  const int m = A.r, n = A.c, k = std::min(m, n);
  for (int i = 0; i < k; ++i) {
    if (m - i <= 1) continue;
    Real ss = 0.0;
    for (int r = i + 1; r < m; ++r) ss += A(r, i) * A(r, i);
    const Real xnorm = std::sqrt(ss);
    if (xnorm == 0.0) continue;
    const Real alpha = A(i, i);
    const Real beta = -std::copysign(std::hypot(alpha, xnorm), alpha);
    const Real tau = (beta - alpha) / beta;
    const Real scal = 1.0 / (alpha - beta);
    for (int r = i + 1; r < m; ++r) A(r, i) *= scal;
    A(i, i) = beta;
    for (int j = i + 1; j < n; ++j) {
      Real w = A(i, j);
      for (int r = i + 1; r < m; ++r) w += A(r, i) * A(r, j);
      w *= tau;
      A(i, j) -= w;
      for (int r = i + 1; r < m; ++r) A(r, j) -= A(r, i) * w;
    }
  }
  Mat R(k, n);
  for (int i = 0; i < k; ++i)
    for (int j = i; j < n; ++j) R(i, j) = A(i, j);
  return R;
}

// util/cholesky_util.py:89-103 sum_of_sqrtm_factors / triu_via_qr.
Mat sum_of_sqrtm_factors(const Mat& R1, const Mat& R2) { return qr_r(vstack(R1, R2)); }

// backend/linalg.py:48-53 solve_triangular (upper / lower), column by column.
Mat solve_triu(const Mat& R, const Mat& B) {
  Mat X = B;
  const int n = R.r;
  for (int col = 0; col < B.c; ++col)
    for (int i = n - 1; i >= 0; --i) {
      Real s = B(i, col);
      for (int j = i + 1; j < n; ++j) s -= R(i, j) * X(j, col);
      X(i, col) = s / R(i, i);
    }
  return X;
}

Mat solve_tril(const Mat& L, const Mat& B) {
  Mat X = B;
  const int n = L.r;
  for (int col = 0; col < B.c; ++col)
    for (int i = 0; i < n; ++i) {
      Real s = B(i, col);
      for (int j = 0; j < i; ++j) s -= L(i, j) * X(j, col);
      X(i, col) = s / L(i, i);
    }
  return X;
}

// backend/linalg.py:56-61 lstsq_svd = jnp.linalg.lstsq(...)[0]: minimum-norm
// least squares with the default cutoff rcond = eps * max(m, n) * s_max.
// One-sided Jacobi SVD (Hestenes) on the square system.
Mat lstsq_svd(const Mat& A, const Mat& B) {
  const int m = A.r, n = A.c;
  Mat U = A;  // columns get orthogonalised
  Mat V = eye(n);
  for (int sweep = 0; sweep < 60; ++sweep) {
    Real off = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        Real alpha = 0, beta = 0, gamma = 0;
        for (int i = 0; i < m; ++i) {
          alpha += U(i, p) * U(i, p);
          beta += U(i, q) * U(i, q);
          gamma += U(i, p) * U(i, q);
        }
        if (gamma == 0.0) continue;
        off = std::max(off, std::fabs(gamma) / std::sqrt(alpha * beta));
        const Real zeta = (beta - alpha) / (2.0 * gamma);
        const Real t = std::copysign(1.0, zeta) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const Real cs = 1.0 / std::sqrt(1.0 + t * t), sn = cs * t;
        for (int i = 0; i < m; ++i) {
          const Real up = U(i, p), uq = U(i, q);
          U(i, p) = cs * up - sn * uq;
          U(i, q) = sn * up + cs * uq;
        }
        for (int i = 0; i < n; ++i) {
          const Real vp = V(i, p), vq = V(i, q);
          V(i, p) = cs * vp - sn * vq;
          V(i, q) = sn * vp + cs * vq;
        }
      }
    if (off < (sizeof(Real) > sizeof(double) ? Real(1e-19) : Real(1e-16))) break;
  }
  Vec s(n);
  Real smax = 0.0;
  for (int j = 0; j < n; ++j) {
    Real ss = 0;
    for (int i = 0; i < m; ++i) ss += U(i, j) * U(i, j);
    s[j] = std::sqrt(ss);
    smax = std::max(smax, s[j]);
  }
  const Real cutoff = 2.220446049250313e-16 * std::max(m, n) * smax;
  Mat X(n, B.c);
  for (int j = 0; j < n; ++j) {
    if (!(s[j] > cutoff)) continue;
    for (int col = 0; col < B.c; ++col) {
      Real ub = 0;
      for (int i = 0; i < m; ++i) ub += (U(i, j) / s[j]) * B(i, col);
      ub /= s[j];
      for (int i = 0; i < n; ++i) X(i, col) += V(i, j) * ub;
    }
  }
  return X;
}

enum { SOLVE_TRIU = 0, SOLVE_LSTSQ = 1 };

Mat solve_upper(const Mat& R, const Mat& B, int mode) {
  return mode == SOLVE_LSTSQ ? lstsq_svd(R, B) : solve_triu(R, B);
}

// util/cholesky_util.py:27-82 revert_conditional.
struct Reverted {
  Mat R_Y, R_XY, G;
};
Reverted revert_conditional(const Mat& R_X_F, const Mat& R_X, const Mat& R_YX, int mode) {
  const int k = R_YX.r, n = R_X.c;
  Mat R(k + n, k + n);
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) R(i, j) = R_YX(i, j);
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < k; ++j) R(k + i, j) = R_X_F(i, j);
    for (int j = 0; j < n; ++j) R(k + i, k + j) = R_X(i, j);
  }
  R = qr_r(R);
  Reverted out;
  out.R_Y = Mat(k, k);
  Mat R12(k, n);
  out.R_XY = Mat(n, n);
  for (int i = 0; i < k; ++i) {
    for (int j = 0; j < k; ++j) out.R_Y(i, j) = R(i, j);
    for (int j = 0; j < n; ++j) R12(i, j) = R(i, k + j);
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) out.R_XY(i, j) = R(k + i, k + j);
  out.G = transpose(solve_upper(out.R_Y, R12, mode));
  return out;
}

// backend/linalg.py solve_lu = jnp.linalg.solve: LU factorisation with
// partial pivoting (the largest |entry| of the column below the diagonal,
// first index on ties, as LAPACK's idamax), then forward and back
// substitution of every right-hand side.
Mat solve_lu(const Mat& M, const Mat& B) {
  const int n = M.r;
  Mat A = M, X = B;
  for (int k = 0; k < n; ++k) {
    int piv = k;
    for (int r = k + 1; r < n; ++r)
      if (std::fabs(A(r, k)) > std::fabs(A(piv, k))) piv = r;
    if (piv != k) {
      for (int j = 0; j < n; ++j) std::swap(A(k, j), A(piv, j));
      for (int j = 0; j < X.c; ++j) std::swap(X(k, j), X(piv, j));
    }
    const Real inv_p = 1.0 / A(k, k);
    for (int r = k + 1; r < n; ++r) {
      const Real l = A(r, k) * inv_p;
      A(r, k) = l;
      for (int j = k + 1; j < n; ++j) A(r, j) -= l * A(k, j);
    }
  }
  for (int j = 0; j < X.c; ++j) {
    for (int i = 1; i < n; ++i) {
      Real s = X(i, j);
      for (int c = 0; c < i; ++c) s -= A(i, c) * X(c, j);
      X(i, j) = s;
    }
    for (int i = n - 1; i >= 0; --i) {
      Real s = X(i, j);
      for (int c = i + 1; c < n; ++c) s -= A(i, c) * X(c, j);
      X(i, j) = s / A(i, i);
    }
  }
  return X;
}

// util/gram_util.py:17-77 exp_gram_cholesky with pade_and_legendre_9
// (:261-331), the float64 choice of ssm_impl_dense.py:611-617: e^A and a
// lower-triangular U with U U^T = int_0^1 e^{sA} B B^T e^{sA^T} ds, by
// scaling A (and B) by 2^-s, a degree-9 Pade approximant of e^A with a
// Legendre-weighted approximation of the Gramian's square root, and s
// doublings, each a QR of [U, e^A U]; columns of U are signed so that its
// diagonal is non-negative.
void exp_gram(const Mat& A_in, const Mat& B_in, Mat& eA, Mat& U) {
  static const Real pade[10] = {17643225600.0, 8821612800.0, 2075673600.0, 302702400.0, 30270240.0,
                                2162160.0,     110880.0,     3960.0,       90.0,        1.0};
  static const Real leg[10][10] = {
      {17643225600.0, 0, 605404800.0, 0, 4324320.0, 0, 7920.0, 0, 2.0, 0},
      {0, 8821612800.0, 0, 155675520.0, 0, 617760.0, 0, 528.0, 0, 0},
      {0, 0, 1470268800.0, 0, 15444000.0, 0, 34320.0, 0, 10.0, 0},
      {0, 0, 0, 147026880.0, 0, 960960.0, 0, 1092.0, 0, 0},
      {0, 0, 0, 0, 10501920.0, 0, 42120.0, 0, 18.0, 0},
      {0, 0, 0, 0, 0, 583440.0, 0, 1320.0, 0, 0},
      {0, 0, 0, 0, 0, 0, 26520.0, 0, 26.0, 0},
      {0, 0, 0, 0, 0, 0, 0, 1020.0, 0, 0},
      {0, 0, 0, 0, 0, 0, 0, 0, 34.0, 0},
      {0, 0, 0, 0, 0, 0, 0, 0, 0, 1.0}};
  const Real eta = 0.41098379928173995, q = 9.0;
  const int n = A_in.r, m = B_in.c;
  // _exp_gram_cholesky_init: the number of doublings
  Real a_l1 = 0.0;
  for (int j = 0; j < n; ++j) {
    Real s = 0.0;
    for (int i = 0; i < n; ++i) s += std::fabs(A_in(i, j));
    a_l1 = std::max(a_l1, s);
  }
  const Real s1 = std::log2(a_l1 / eta), s2 = std::log2((Real)(n - 1) / q);
  Real num = std::max((Real)0.0, std::ceil(std::max(s1, s2)));
  if (!(num <= 64)) num = 64;  // bounded (as the device): a larger norm gives a non-finite transition
  const Real two_num = std::pow((Real)2.0, num);
  const Mat A = scale(1.0 / two_num, A_in);
  Mat B = B_in;
  const Real sq = std::sqrt(two_num);
  for (auto& x : B.a) x = x / sq;
  // pade_and_legendre_9.init
  const Mat I = eye(n);
  const Mat A2 = matmul(A, A), A4 = matmul(A2, A2), A6 = matmul(A4, A2), A8 = matmul(A6, A2);
  const Mat Up = matmul(A, add(add(add(add(scale(pade[9], A8), scale(pade[7], A6)), scale(pade[5], A4)),
                                   scale(pade[3], A2)),
                               scale(pade[1], I)));
  const Mat V = add(add(add(add(scale(pade[8], A8), scale(pade[6], A6)), scale(pade[4], A4)), scale(pade[2], A2)),
                    scale(pade[0], I));
  const Mat VmU = sub(V, Up);
  eA = solve_lu(VmU, add(V, Up));
  Mat P = A2;
  Mat PB = matmul(P, B);
  std::vector<Mat> Le(5, Mat(n, m)), Lo(5, Mat(n, m));
  Le[0] = add(scale(leg[0][2], PB), scale(leg[0][0], B));
  Le[1] = scale(leg[2][2], PB);
  Lo[0] = add(scale(leg[1][3], PB), scale(leg[1][1], B));
  Lo[1] = scale(leg[3][3], PB);
  for (int k = 2; k <= 4; ++k) {
    P = matmul(A2, P);
    PB = matmul(P, B);
    for (int i = 0; i < 5; ++i) {
      Le[i] = add(Le[i], scale(leg[2 * i][2 * k], PB));
      Lo[i] = add(Lo[i], scale(leg[2 * i + 1][2 * k + 1], PB));
    }
  }
  for (int i = 0; i < 5; ++i) Lo[i] = matmul(A, Lo[i]);
  Mat rhs(n, 10 * m);
  for (int i = 0; i < 10; ++i) {
    const Mat& ell = (i % 2 == 0) ? Le[i / 2] : Lo[i / 2];
    const Real nrm = std::sqrt((Real)(2 * i + 1));
    for (int r = 0; r < n; ++r)
      for (int c = 0; c < m; ++c) rhs(r, i * m + c) = ell(r, c) / nrm;
  }
  const Mat Ls = solve_lu(VmU, rhs);
  const Mat Rt = qr_r(transpose(Ls));  // min(10m, n) x n
  U = Mat(n, n);
  for (int i = 0; i < Rt.r; ++i)
    for (int j = 0; j < n; ++j) U(j, i) = Rt(i, j);
  // _exp_gram_cholesky_double
  for (int it = 0; it < (int)num; ++it) {
    const Mat eU = matmul(eA, U);
    Mat stack(n, 2 * n);
    for (int r = 0; r < n; ++r)
      for (int c = 0; c < n; ++c) {
        stack(r, c) = U(r, c);
        stack(r, n + c) = eU(r, c);
      }
    U = transpose(qr_r(transpose(stack)));
    eA = matmul(eA, eA);
  }
  for (int j = 0; j < n; ++j) {
    Real sgn = U(j, j) > 0 ? 1.0 : (U(j, j) < 0 ? -1.0 : 0.0);
    if (sgn == 0.0) sgn = 1.0;
    for (int i = 0; i < n; ++i) U(i, j) = U(i, j) * sgn;
  }
}

// util/cholesky_util.py:106-176 cholesky_hilbert (Kahan's recurrence).
Mat cholesky_hilbert(int n) {
  const Real Kf = 0.0;
  Vec dr(n), f(n, 1.0 + Kf);
  for (int i = 0; i < n; ++i) dr[i] = std::sqrt((Real)(2 * i + 1));
  for (int idx = 1; idx < n; ++idx) {
    const Real prev = f[idx - 1], idxf = idx;
    f[idx] = (((prev / idxf) * (Kf + 2.0 * idxf)) / (Kf + idxf)) * (Kf + 2.0 * idxf + 1.0);
  }
  for (auto& x : f) x = 1.0 / x;
  Mat U = eye(n);
  for (int j = 1; j < n; ++j)
    for (int kk = 0; kk < j; ++kk) {
      const int i = j - 1 - kk;
      const Real denom = (Real)(j - i), factor = Kf + (Real)(i + 1) + (Real)(j + 1);
      U(i, j) = (U(i + 1, j) / denom) * factor;
    }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) U(i, j) = U(i, j) * (dr[i] * f[j]);
  Mat L = transpose(U);
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) L(i, j) = 0.0;
  return L;
}

Real factorial_int(int k) {
  Real f = 1.0;
  for (int i = 2; i <= k; ++i) f *= i;
  return f;
}

Real binom_int(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  return factorial_int(n) / (factorial_int(n - k) * factorial_int(k));
}

// _probdiffeq/utilities.py:57-71 system_matrices_1d_iwp.
void system_matrices_1d_iwp(int nu, Mat& A1, Mat& Q1) {
  const int N = nu + 1;
  A1 = Mat(N, N);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) A1(i, j) = binom_int(nu - i, nu - j);
  Mat L = cholesky_hilbert(N);
  Mat Lf(N, N);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) Lf(i, j) = L(N - 1 - i, j);
  Mat Q = transpose(qr_r(transpose(Lf)));
  for (int j = 0; j < N; ++j) {
    Real s = (Q(j, j) > 0) ? 1.0 : (Q(j, j) < 0 ? -1.0 : 1.0);
    for (int i = 0; i < N; ++i) Q(i, j) *= s;
  }
  Q1 = Q;
}

// _probdiffeq/utilities.py:74-84 preconditioner_taylor.  Factorials are
// exact integers here (the reference forms exp(lgamma(k+1))).
void preconditioner(int nu, Real dt, Vec& p, Vec& pinv) {
  const int N = nu + 1;
  p.assign(N, 0.0);
  pinv.assign(N, 0.0);
  for (int i = 0; i < N; ++i) {
    const int pw = nu - i;
    const Real s = factorial_int(pw);
    p[i] = std::pow(dt, (Real)pw) / s;
    pinv[i] = std::pow(dt, -(Real)pw) * s;
  }
}

// ------------------------------------------------------- vector fields
enum VfId { VF_LOTKA_VOLTERRA = 0, VF_VAN_DER_POL = 1, VF_BURGERS = 2, VF_LINEAR = 3, VF_LOGISTIC = 4 };

struct Vf {
  int id = 0, d = 0;
  const Real* prm = nullptr;
  // f(y, t) -> out (d)
  void eval(const Real* y, Real t, Real* out) const {
    switch (id) {
      case VF_LOTKA_VOLTERRA: {
        const Real a = prm[0], b = prm[1], c = prm[2], dd = prm[3];
        out[0] = a * y[0] - b * y[0] * y[1];
        out[1] = -c * y[1] + dd * y[0] * y[1];
        break;
      }
      case VF_VAN_DER_POL: {
        const Real mu = prm[0];
        out[0] = y[1];
        out[1] = mu * ((1.0 - y[0] * y[0]) * y[1]) - y[0];
        break;
      }
      case VF_BURGERS: {
        const Real visc = prm[0], dx = prm[1];
        for (int j = 0; j < d; ++j) {
          const Real um = y[(j + d - 1) % d], u0 = y[j], up = y[(j + 1) % d];
          out[j] = -u0 * ((up - um) / (2.0 * dx)) + visc * ((up - 2.0 * u0 + um) / (dx * dx));
        }
        break;
      }
      case VF_LINEAR: {
        for (int j = 0; j < d; ++j) out[j] = prm[j] * y[j] + t;
        break;
      }
      case VF_LOGISTIC: {
        const Real r = prm[0];
        for (int j = 0; j < d; ++j) out[j] = r * y[j] * (1.0 - y[j]);
        break;
      }
    }
  }
  // Jacobian df/dy (d x d, row-major).
  void jac(const Real* y, Real t, Real* J) const {
    (void)t;
    std::fill(J, J + d * d, 0.0);
    switch (id) {
      case VF_LOTKA_VOLTERRA: {
        const Real a = prm[0], b = prm[1], c = prm[2], dd = prm[3];
        J[0] = a - b * y[1];
        J[1] = -b * y[0];
        J[2] = dd * y[1];
        J[3] = -c + dd * y[0];
        break;
      }
      case VF_VAN_DER_POL: {
        const Real mu = prm[0];
        J[0] = 0.0;
        J[1] = 1.0;
        J[2] = mu * (-2.0 * y[0] * y[1]) - 1.0;
        J[3] = mu * (1.0 - y[0] * y[0]);
        break;
      }
      case VF_BURGERS: {
        const Real visc = prm[0], dx = prm[1];
        for (int j = 0; j < d; ++j) {
          const Real um = y[(j + d - 1) % d], u0 = y[j], up = y[(j + 1) % d];
          const int jm = (j + d - 1) % d, jp = (j + 1) % d;
          J[j * d + j] += -((up - um) / (2.0 * dx)) - 2.0 * visc / (dx * dx);
          J[j * d + jp] += -u0 / (2.0 * dx) + visc / (dx * dx);
          J[j * d + jm] += u0 / (2.0 * dx) + visc / (dx * dx);
        }
        break;
      }
      case VF_LINEAR: {
        for (int j = 0; j < d; ++j) J[j * d + j] = prm[j];
        break;
      }
      case VF_LOGISTIC: {
        const Real r = prm[0];
        for (int j = 0; j < d; ++j) J[j * d + j] = r * (1.0 - 2.0 * y[j]);
        break;
      }
    }
  }
};

// Taylor-mode arithmetic for the initial Taylor coefficients (restates
// jet_expansion_algorithms.py:49-103 / :155-177 for the registered vector
// fields).  c[k][j] are normalised Taylor coefficients; the returned
// derivatives are k! c[k].
void taylor_coefficients(const Vf& vf, const Real* y0, Real t0, int num, Real* out) {
  const int d = vf.d, K = num + 1;
  std::vector<Vec> c(K, Vec(d, 0.0));
  for (int j = 0; j < d; ++j) c[0][j] = y0[j];
  auto cauchy = [&](const Vec* X, const Vec* Y, int k, int j, int jx, int jy) {
    Real s = 0.0;
    for (int i = 0; i <= k; ++i) s += X[i][jx] * Y[k - i][jy];
    (void)j;
    return s;
  };
  for (int k = 0; k < num; ++k) {
    // coefficient k of f(y(t0 + s))
    Vec fk(d, 0.0);
    switch (vf.id) {
      case VF_LOTKA_VOLTERRA: {
        const Real a = vf.prm[0], b = vf.prm[1], cc = vf.prm[2], dd = vf.prm[3];
        const Real y0y1 = cauchy(c.data(), c.data(), k, 0, 0, 1);
        fk[0] = a * c[k][0] - b * y0y1;
        fk[1] = -cc * c[k][1] + dd * y0y1;
        break;
      }
      case VF_VAN_DER_POL: {
        const Real mu = vf.prm[0];
        // (1 - y0^2) y1 coefficient k
        std::vector<Vec> sq(k + 1, Vec(1, 0.0));
        for (int m = 0; m <= k; ++m) {
          Real s = 0.0;
          for (int i = 0; i <= m; ++i) s += c[i][0] * c[m - i][0];
          sq[m][0] = (m == 0 ? 1.0 : 0.0) - s;
        }
        Real s = 0.0;
        for (int i = 0; i <= k; ++i) s += sq[i][0] * c[k - i][1];
        fk[0] = c[k][1];
        fk[1] = mu * s - c[k][0];
        break;
      }
      case VF_BURGERS: {
        const Real visc = vf.prm[0], dx = vf.prm[1];
        for (int j = 0; j < d; ++j) {
          const int jm = (j + d - 1) % d, jp = (j + 1) % d;
          Real s = 0.0;
          for (int i = 0; i <= k; ++i) s += c[i][j] * ((c[k - i][jp] - c[k - i][jm]) / (2.0 * dx));
          fk[j] = -s + visc * ((c[k][jp] - 2.0 * c[k][j] + c[k][jm]) / (dx * dx));
        }
        break;
      }
      case VF_LINEAR: {
        for (int j = 0; j < d; ++j) fk[j] = vf.prm[j] * c[k][j] + (k == 0 ? t0 : (k == 1 ? 1.0 : 0.0));
        break;
      }
      case VF_LOGISTIC: {
        const Real r = vf.prm[0];
        for (int j = 0; j < d; ++j) {
          Real s = 0.0;
          for (int i = 0; i <= k; ++i) s += c[i][j] * c[k - i][j];
          fk[j] = r * (c[k][j] - s);
        }
        break;
      }
    }
    for (int j = 0; j < d; ++j) c[k + 1][j] = fk[j] / (Real)(k + 1);
  }
  for (int k = 0; k < K; ++k)
    for (int j = 0; j < d; ++j) out[k * d + j] = factorial_int(k) * c[k][j];
}

// ------------------------------------------------------ configuration
enum { FACT_DENSE = 0, FACT_ISO = 1, FACT_BLOCKDIAG = 2 };
enum { CONSTRAINT_TS0 = 0, CONSTRAINT_TS1 = 1 };
enum { STRAT_FILTER = 0, STRAT_FIXEDPOINT = 1, STRAT_FIXEDINTERVAL = 2 };
enum { SOLVER_PLAIN = 0, SOLVER_DYNAMIC = 1, SOLVER_MLE = 2 };
enum { NORM_SCALE_THEN_RMS = 0, NORM_RMS_THEN_SCALE = 1 };
enum { CONTROL_INTEGRAL = 0, CONTROL_PI = 1 };

}  // namespace

extern "C" {
struct pdqo_cfg {
  int fact, nu, d, constraint, strategy, solver, error_norm, control, vf_id, clip_dt;
  Real rtol, atol, dt0, eps, damp;
  Real safety, factor_min, factor_max, exponent_integral, exponent_proportional;
  Real base_scale;  // prior output scale (scalar, broadcast)
  int correct_asymptotic_underconfidence, max_attempts;
  int re_linearize;  // solver_dynamic(re_linearize_after_calibration=True)
  // error estimator: 0 error_residual_std, 1 error_state_std; options
  int error_estimator, error_re_linearize, error_derivative_idx, error_per_unit_step;
  // prior: 0 integrated Wiener process, 1 exponential (dense only) whose
  // linear drift's bottom block (d x (nu+1)d, row-major, derivative-major
  // columns) is prior_drift
  int prior;
  const Real* prior_drift;
  // loss_lml_timeseries(tcoeff_index=): the Taylor coefficient the data observe
  int tcoeff_index;
  // from_mean_and_std with a std per covariance row (element-wise is_exact,
  // diffuse_derivatives; ssm_impl_isotropic.py:409-447): n values, or NULL
  // for the scalar init_std on every row
  const Real* init_std_rows;
  // loss_lml_timeseries(solve_triu=): 0 linalg.lstsq_svd (the default), 1
  // linalg.solve_triu
  int lml_solve;
  // ts1 Jacobian handler: 0 jacobian_materialize (exact), 1
  // jacobian_monte_carlo_rev, 2 jacobian_monte_carlo_fwd; the number of
  // Rademacher probes and the seed of their stream
  int jacobian, jac_probes;
  uint32_t jac_seed;
};
}

namespace {

// ============================================================== SSM ops
// Dense (ssm_impl_dense.py) and isotropic (ssm_impl_isotropic.py) share the
// representation Normal{m (rows x cols), L (rows x rows)}; dense means are
// columns (n x 1) in derivative-major order, isotropic means are N x d.
struct Normal {
  Mat m, L;
};
struct Cond {
  Mat A;
  Normal noise;
  Vec tl, to;  // to_latent (A.c), to_observed (A.r)
};

Vec ones(int n) { return Vec(n, 1.0); }

Vec inv(const Vec& v) {
  Vec r(v.size());
  for (size_t i = 0; i < v.size(); ++i) r[i] = 1.0 / v[i];
  return r;
}

struct SSMBase {
  int fact;
  bool iso() const { return fact == FACT_ISO; }

  // ssm_impl_dense.py:18-22 / ssm_impl_isotropic.py:75-79 apply_flat.
  Normal apply_flat(const Cond& C, const Mat& x) const {
    Mat xl = rowscale(C.tl, x);
    Normal out;
    if (iso()) {
      Mat toA = rowscale(C.to, C.A);
      out.m = add(matmul(toA, xl), C.noise.m);
    } else {
      out.m = rowscale(C.to, add(matmul(C.A, xl), C.noise.m));
    }
    out.L = rowscale(C.to, C.noise.L, true);
    return out;
  }

  // ssm_impl_dense.py:24-33 / ssm_impl_isotropic.py:81-89 (blockdiag:
  // ssm_impl_blockdiag.py:28-43 takes |to_latent| on the factor).
  Normal marginalise(const Cond& C, const Normal& rv, bool abs_tl = false) const {
    Mat mean = rowscale(C.tl, rv.m);
    Mat chol = rowscale(C.tl, rv.L, abs_tl);
    Mat R = sum_of_sqrtm_factors(transpose(matmul(C.A, chol)), transpose(C.noise.L));
    Normal out;
    out.m = rowscale(C.to, add(matmul(C.A, mean), C.noise.m));
    out.L = rowscale(C.to, transpose(R), true);
    return out;
  }

  // ssm_impl_dense.py:35-49 / ssm_impl_isotropic.py:91-105 merge.
  Cond merge(const Cond& self, const Cond& other) const {
    Vec T(self.tl.size());
    for (size_t i = 0; i < T.size(); ++i) T[i] = self.tl[i] * other.to[i];
    Cond out;
    out.A = matmul(self.A, rowscale(T, other.A));
    out.noise.m = add(matmul(self.A, rowscale(T, other.noise.m)), self.noise.m);
    Mat R1 = transpose(matmul(self.A, rowscale(T, other.noise.L, true)));
    Mat R2 = transpose(self.noise.L);
    out.noise.L = transpose(sum_of_sqrtm_factors(R1, R2));
    out.tl = other.tl;
    out.to = self.to;
    return out;
  }

  // ssm_impl_dense.py:51-77 / ssm_impl_isotropic.py:107-133 revert.
  void revert(const Cond& C, const Normal& rv, int mode, Normal& observed, Cond& bw) const {
    Mat mean = rowscale(C.tl, rv.m);
    Mat chol = rowscale(C.tl, rv.L, true);
    Mat R_X_F = transpose(matmul(C.A, chol));
    Mat R_X = transpose(chol);
    Mat R_YX = transpose(C.noise.L);
    Reverted rc = revert_conditional(R_X_F, R_X, R_YX, mode);
    Mat mean_observed = add(matmul(C.A, mean), C.noise.m);
    Mat mean_corrected = sub(mean, matmul(rc.G, mean_observed));
    bw.A = rc.G;
    bw.noise.m = mean_corrected;
    bw.noise.L = transpose(rc.R_XY);
    bw.tl = inv(C.to);
    bw.to = inv(C.tl);
    observed.m = rowscale(C.to, mean_observed);
    observed.L = rowscale(C.to, transpose(rc.R_Y), true);
  }

  // ssm_impl_dense.py:156-160 / ssm_impl_isotropic.py:203-207.
  Real residual_whitened_rms_zero(const Normal& rv) const {
    Mat dx = rv.m;
    if (!iso())
      for (auto& x : dx.a) x = 0.0 - x;
    Mat w = solve_tril(rv.L, dx);
    return fro(w) / std::sqrt((Real)rv.m.size());
  }

  // ssm_impl_dense.py:141-150 (row norms via per-row qr_r) /
  // ssm_impl_isotropic.py:189-197 (vector_norm of rows).
  Vec std_rows(const Normal& rv) const {
    const int n = rv.L.r;
    Vec s(n);
    for (int i = 0; i < n; ++i) {
      if (iso()) {
        Real ss = 0;
        for (int j = 0; j < rv.L.c; ++j) ss += rv.L(i, j) * rv.L(i, j);
        s[i] = std::sqrt(ss);
      } else {
        Mat col(rv.L.c, 1);
        for (int j = 0; j < rv.L.c; ++j) col(j, 0) = rv.L(i, j);
        s[i] = std::fabs(qr_r(col)(0, 0));
      }
    }
    return s;
  }

  // ssm_impl_dense.py:170-180 / ssm_impl_isotropic.py:217-236 logpdf.
  Real logpdf(const Normal& rv, const Mat& u) const {
    Mat chol = transpose(qr_r(transpose(rv.L)));
    Real slogdet = 0.0;
    for (int i = 0; i < chol.r; ++i) slogdet += std::log(std::fabs(chol(i, i)));
    const Real c2pi = std::log(Real(2) * Real(3.14159265358979323846264338327950288L));
    if (!iso()) {
      Mat dx = sub(u, rv.m);
      Mat w = solve_tril(chol, dx);
      Real sq = 0;
      for (Real x : w.a) sq += x * x;
      return -1.0 / 2.0 * sq - (Real)u.size() / 2.0 * c2pi - slogdet;
    }
    Real total = 0.0;
    for (int col = 0; col < u.c; ++col) {
      Mat dx(u.r, 1);
      for (int i = 0; i < u.r; ++i) dx(i, 0) = u(i, col) - rv.m(i, col);
      Mat w = solve_tril(chol, dx);
      Real maha = 0;
      for (Real x : w.a) maha += x * x;
      const Real logdet_term = 2.0 * slogdet;
      total += -0.5 * (logdet_term + maha + (Real)u.r * c2pi);
    }
    return total;
  }

  // ssm_impl_dense.py:196-200 / ssm_impl_isotropic.py:260-266.
  Cond identity_conditional(const Normal& like) const {
    const int n = like.L.r;
    Cond C;
    C.A = eye(n);
    C.noise.m = Mat(like.m.r, like.m.c);
    C.noise.L = Mat(n, n);
    C.tl = ones(n);
    C.to = ones(n);
    return C;
  }
};

// ========================================================= problem state
struct Problem {
  pdqo_cfg cfg;
  int N, d, n;   // n = covariance dimension (dense: N*d; iso, blockdiag: N)
  int nblk;      // blockdiag: d blocks; otherwise 1
  Mat A1, Q1;    // 1-d IWP matrices
  Mat A, Q;      // dense: kron(A1, I), kron(Q1, I); iso/bd: A1, Q1
  Mat Aexp, Bexp;  // exponential prior: drift (n x n) and diffusion (n x d)
  Vf vf;
  SSMBase ops;

  void setup(const pdqo_cfg& c, const Real* params) {
    cfg = c;
    N = c.nu + 1;
    d = c.d;
    system_matrices_1d_iwp(c.nu, A1, Q1);
    vf.id = c.vf_id;
    vf.d = d;
    vf.prm = params;
    ops.fact = c.fact == FACT_ISO ? FACT_ISO : FACT_DENSE;
    if (c.fact == FACT_DENSE) {
      n = N * d;
      nblk = 1;
      A = Mat(n, n);
      Q = Mat(n, n);
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
          for (int k = 0; k < d; ++k) {
            A(i * d + k, j * d + k) = A1(i, j);
            Q(i * d + k, j * d + k) = Q1(i, j) * c.base_scale;  // kron(Q1, Lambda)
          }
      if (c.prior == 1) {
        // prior_exponential_diffuse (ssm_impl_dense.py:541-629): A = kron(a, I)
        // with a the shift matrix and the drift as its last d rows;
        // B = kron(e_last, Lambda)
        Aexp = Mat(n, n);
        for (int i = 0; i + 1 < N; ++i)
          for (int k = 0; k < d; ++k) Aexp(i * d + k, (i + 1) * d + k) = 1.0;
        for (int r = 0; r < d; ++r)
          for (int j = 0; j < n; ++j) Aexp((N - 1) * d + r, j) = c.prior_drift[r * n + j];
        Bexp = Mat(n, d);
        for (int k = 0; k < d; ++k) Bexp((N - 1) * d + k, k) = c.base_scale;
      }
    } else {
      n = N;
      nblk = c.fact == FACT_BLOCKDIAG ? d : 1;
      A = A1;
      Q = Q1;
    }
  }
  bool bd() const { return cfg.fact == FACT_BLOCKDIAG; }
};

// A "state" bundles per-block normals (one block unless blockdiag).
struct BNormal {
  std::vector<Normal> b;
};
struct BCond {
  std::vector<Cond> b;
};

// *WienerIntegrated.transition (ssm_impl_dense.py:347-363,
// ssm_impl_isotropic.py:367-378, ssm_impl_blockdiag.py:406-432).
BCond transition(const Problem& P, Real dt, const Vec& os) {
  Vec p1, pinv1;
  preconditioner(P.cfg.nu, dt, p1, pinv1);
  BCond out;
  out.b.resize(P.nblk);
  for (int blk = 0; blk < P.nblk; ++blk) {
    Cond& C = out.b[blk];
    if (P.cfg.fact == FACT_DENSE) {
      C.tl.resize(P.n);
      C.to.resize(P.n);
      for (int i = 0; i < P.N; ++i)
        for (int k = 0; k < P.d; ++k) {
          C.to[i * P.d + k] = p1[i];
          C.tl[i * P.d + k] = pinv1[i];
        }
      C.noise.m = Mat(P.n, 1);
      if (P.cfg.prior == 1) {
        // DenseExponential.transition (ssm_impl_dense.py:405-423)
        Mat Ap(P.n, P.n), Bp(P.n, P.d);
        for (int i = 0; i < P.n; ++i)
          for (int j = 0; j < P.n; ++j) Ap(i, j) = ((dt * C.tl[i]) * P.Aexp(i, j)) * C.to[j];
        for (int i = 0; i < P.n; ++i)
          for (int k = 0; k < P.d; ++k) Bp(i, k) = (std::sqrt(std::fabs(dt)) * std::fabs(C.tl[i])) * P.Bexp(i, k);
        Mat eA, U;
        exp_gram(Ap, Bp, eA, U);
        C.A = eA;
        C.noise.L = scale(os[0], U);
      } else {
        C.A = P.A;
        C.noise.L = scale(std::sqrt(std::fabs(dt)) * os[0], P.Q);
      }
    } else if (P.cfg.fact == FACT_ISO) {
      C.to = p1;
      C.tl = pinv1;
      C.A = P.A;
      C.noise.m = Mat(P.N, P.d);
      const Real s = std::sqrt(std::fabs(dt)) * P.cfg.base_scale * os[0];
      C.noise.L = scale(s, P.Q);
    } else {
      C.to = p1;
      C.tl = pinv1;
      C.A = P.A;
      C.noise.m = Mat(P.N, 1);
      const Real s = std::sqrt(std::fabs(dt)) * (P.cfg.base_scale * os[blk]);
      C.noise.L = scale(s, P.Q);
    }
  }
  return out;
}

// Block-wise wrappers.
BNormal apply_flat(const Problem& P, const BCond& C, const BNormal& x) {
  BNormal out;
  out.b.resize(P.nblk);
  for (int i = 0; i < P.nblk; ++i) out.b[i] = P.ops.apply_flat(C.b[i], x.b[i].m);
  return out;
}
BNormal marginalise(const Problem& P, const BCond& C, const BNormal& rv) {
  BNormal out;
  out.b.resize(P.nblk);
  for (int i = 0; i < P.nblk; ++i) out.b[i] = P.ops.marginalise(C.b[i], rv.b[i], P.bd());
  return out;
}
BCond merge(const Problem& P, const BCond& self, const BCond& other) {
  BCond out;
  out.b.resize(P.nblk);
  for (int i = 0; i < P.nblk; ++i) out.b[i] = P.ops.merge(self.b[i], other.b[i]);
  return out;
}
void revert(const Problem& P, const BCond& C, const BNormal& rv, int mode, BNormal& obs, BCond& bw) {
  obs.b.resize(P.nblk);
  bw.b.resize(P.nblk);
  for (int i = 0; i < P.nblk; ++i) P.ops.revert(C.b[i], rv.b[i], mode, obs.b[i], bw.b[i]);
}
BCond identity_conditional(const Problem& P, const BNormal& like) {
  BCond out;
  out.b.resize(P.nblk);
  for (int i = 0; i < P.nblk; ++i) out.b[i] = P.ops.identity_conditional(like.b[i]);
  return out;
}
// residual_whitened_rms against zero data: scalar (dense/iso) or per block.
Vec residual_whitened_rms_zero(const Problem& P, const BNormal& rv) {
  Vec out(P.nblk);
  for (int i = 0; i < P.nblk; ++i) out[i] = P.ops.residual_whitened_rms_zero(rv.b[i]);
  return out;
}
void rescale_cholesky(const Problem& P, BNormal& rv, const Vec& f) {
  for (int i = 0; i < P.nblk; ++i)
    for (auto& x : rv.b[i].L.a) x = f[P.bd() ? i : 0] * x;
}
void rescale_noise(const Problem& P, BCond& c, const Vec& f) {
  for (int i = 0; i < P.nblk; ++i)
    for (auto& x : c.b[i].noise.L.a) x = f[P.bd() ? i : 0] * x;
}
// The zeroth Taylor coefficient (u(t)) as a flat vector of length d.
Vec u0_of(const Problem& P, const BNormal& rv) {
  Vec u(P.d);
  if (P.cfg.fact == FACT_DENSE)
    for (int k = 0; k < P.d; ++k) u[k] = rv.b[0].m(k, 0);
  else if (P.cfg.fact == FACT_ISO)
    for (int k = 0; k < P.d; ++k) u[k] = rv.b[0].m(0, k);
  else
    for (int k = 0; k < P.d; ++k) u[k] = rv.b[k].m(0, 0);
  return u;
}
Vec coeff_of(const Problem& P, const BNormal& rv, int i) {
  Vec u(P.d);
  if (P.cfg.fact == FACT_DENSE)
    for (int k = 0; k < P.d; ++k) u[k] = rv.b[0].m(i * P.d + k, 0);
  else if (P.cfg.fact == FACT_ISO)
    for (int k = 0; k < P.d; ++k) u[k] = rv.b[0].m(i, k);
  else
    for (int k = 0; k < P.d; ++k) u[k] = rv.b[k].m(i, 0);
  return u;
}

// ------------------------------------------- Monte Carlo Jacobian handlers
enum { JAC_EXACT = 0, JAC_MC_REV = 1, JAC_MC_FWD = 2 };

// One Rademacher entry: the sign of the top bit of the SplitMix64 finaliser
// (Steele, Lea and Flood, OOPSLA 2014) of a counter mixing the seed, the
// probe set, the probe and the coordinate.  The reference draws its probes
// with random.rademacher from a threefry key split at every linearisation
// (jacobians.py:236-238, 255-257); this stream replaces that key chain with
// one probe set per linearisation.
Real rademacher(uint32_t seed, uint32_t set, uint32_t probe, uint32_t coord) {
  uint64_t z = (uint64_t)seed * 0x9E3779B97F4A7C15ull + (uint64_t)set * 0xD1B54A32D192ED03ull +
               (uint64_t)probe * 0x8CB92BA72F3D8DD7ull + (uint64_t)coord * 0xABC98388FB8FAC03ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (z >> 63) ? -1.0 : 1.0;
}

// calculate_diagonal_along_d for one output row (jacobians.py:250-271 in
// reverse mode: vjp(v) * v; :171-192 in forward mode: jvp(v) * v), averaged
// over the probes, with the dense Jacobian Jf (row-major d x d).  Summed
// over k it is calculate_trace_along_d (:230-248, :151-169).
Real mc_diag(const Real* Jf, int d, int k, uint32_t seed, uint32_t set, int probes, bool fwd) {
  Real acc = 0.0;
  for (int s = 0; s < probes; ++s) {
    Real w = 0.0;  // (v^T J)_k or (J v)_k
    for (int j = 0; j < d; ++j) w += rademacher(seed, set, (uint32_t)s, (uint32_t)j) * (fwd ? Jf[k * d + j] : Jf[j * d + k]);
    acc += w * rademacher(seed, set, (uint32_t)s, (uint32_t)k);
  }
  return acc / (Real)probes;
}

// The Jacobian diagonal entry k that the configured handler reports.
Real jac_diag(const pdqo_cfg& c, const Real* Jf, int d, int k, uint32_t set) {
  if (c.jacobian == JAC_EXACT) return Jf[k * d + k];
  return mc_diag(Jf, d, k, c.jac_seed, set, c.jac_probes, c.jacobian == JAC_MC_FWD);
}

// Probe sets a step uses: solver_dynamic with re_linearize_after_calibration
// linearises twice (solvers.py:561-582), every other step once.
uint32_t lin_sets_per_step(const pdqo_cfg& c) {
  return (c.solver == SOLVER_DYNAMIC && c.re_linearize) ? 2u : 1u;
}

// Linearisations.  ts0: *OdeTs0.linearize (ssm_impl_dense.py:237-259,
// ssm_impl_isotropic.py:298-317, ssm_impl_blockdiag.py:123-144).
// ts1: DenseResidual.linearize (ssm_impl_dense.py:262-334) with
// residual_from_ode (problems.py:518-532) and the prior-mean Taylor point.
// `set` is the probe set of the Monte Carlo handlers (isotropic traces and
// blockdiag diagonals; the dense factorisation materialises the Jacobian,
// jacobians.py:221-228, and draws none).
BCond linearize(const Problem& P, const BNormal& rv, Real t, uint32_t set = 0) {
  const int d = P.d, N = P.N;
  Vec y0 = u0_of(P, rv);
  Vec f(d);
  P.vf.eval(y0.data(), t, f.data());
  BCond out;
  out.b.resize(P.nblk);
  const Real damp = P.cfg.damp;
  if (P.cfg.fact == FACT_DENSE) {
    Cond& C = out.b[0];
    const int n = P.n;
    if (P.cfg.constraint == CONSTRAINT_TS0) {
      C.A = Mat(d, n);
      for (int k = 0; k < d; ++k) C.A(k, 1 * d + k) = 1.0;
      C.noise.m = Mat(d, 1);
      for (int k = 0; k < d; ++k) C.noise.m(k, 0) = -f[k];
    } else {
      // r(m) = m_1 - f(m_0); J = [-J_f, I, 0...]; fx = r(xi) - J xi.
      Vec Jf(d * d);
      P.vf.jac(y0.data(), t, Jf.data());
      C.A = Mat(d, n);
      for (int i = 0; i < d; ++i) {
        for (int j = 0; j < d; ++j) C.A(i, j) = -Jf[i * d + j];
        C.A(i, d + i) = 1.0;
      }
      Vec y1 = coeff_of(P, rv, 1);
      Mat Jxi = matmul(C.A, rv.b[0].m);
      C.noise.m = Mat(d, 1);
      for (int k = 0; k < d; ++k) C.noise.m(k, 0) = (y1[k] - f[k]) - Jxi(k, 0);
    }
    C.noise.L = Mat(d, d);
    for (int k = 0; k < d; ++k) C.noise.L(k, k) = damp;
    C.tl = ones(n);
    C.to = ones(d);
  } else if (P.cfg.fact == FACT_ISO) {
    Cond& C = out.b[0];
    C.A = Mat(1, N);
    C.A(0, 1) = 1.0;
    C.noise.m = Mat(1, d);
    if (P.cfg.constraint == CONSTRAINT_TS0) {
      for (int k = 0; k < d; ++k) C.noise.m(0, k) = -f[k];
    } else {
      // IsotropicResidual.linearize (ssm_impl_isotropic.py:321-355) with the
      // exact trace of jacobian_materialize (jacobians.py:100-106): for
      // r = m_1 - f(m_0) the trace over d of dr/dm is [-tr J_f, d, 0...],
      // divided by d; fx = r(m) - linop m.
      Vec Jf(d * d);
      P.vf.jac(y0.data(), t, Jf.data());
      Real tr = 0.0;
      for (int k = 0; k < d; ++k) tr += -jac_diag(P.cfg, Jf.data(), d, k, set);
      C.A(0, 0) = tr / (Real)d;
      C.A(0, 1) = (Real)d / (Real)d;
      const Mat& m = rv.b[0].m;
      for (int k = 0; k < d; ++k) {
        Real lin = 0.0;
        for (int n = 0; n < N; ++n) lin += C.A(0, n) * m(n, k);
        C.noise.m(0, k) = (m(1, k) - f[k]) - lin;
      }
    }
    C.noise.L = Mat(1, 1);
    C.noise.L(0, 0) = damp;
    C.tl = ones(N);
    C.to = ones(1);
  } else {
    // ts0: BlockDiagOdeTs0.linearize (ssm_impl_blockdiag.py:123-144).
    // ts1: BlockDiagResidual.linearize (:147-183) with jacobian_materialize's
    // calculate_diagonal_along_d (jacobians.py:108-113): for r = m_1 - f(m_0)
    // block k's operator is [d r_k / d m_0k, d r_k / d m_1k, 0...] =
    // [-J_f[k][k], 1, 0...] and fx = r_k(m) - operator . m_k.
    Vec Jf;
    if (P.cfg.constraint != CONSTRAINT_TS0) {
      Jf.resize((size_t)d * d);
      P.vf.jac(y0.data(), t, Jf.data());
    }
    for (int k = 0; k < d; ++k) {
      Cond& C = out.b[k];
      C.A = Mat(1, N);
      C.noise.m = Mat(1, 1);
      if (P.cfg.constraint == CONSTRAINT_TS0) {
        C.A(0, 1) = 1.0;
        C.noise.m(0, 0) = -f[k];
      } else {
        C.A(0, 0) = -jac_diag(P.cfg, Jf.data(), d, k, set);
        C.A(0, 1) = 1.0;
        const Mat& m = rv.b[k].m;  // (N, 1)
        const Real r = m(1, 0) - f[k];
        Real lin = 0.0;
        for (int n = 0; n < N; ++n) lin += C.A(0, n) * m(n, 0);
        C.noise.m(0, 0) = r - lin;
      }
      C.noise.L = Mat(1, 1);
      C.noise.L(0, 0) = damp;
      C.tl = ones(N);
      C.to = ones(1);
    }
  }
  return out;
}

// Error-estimate standard deviations of the observed marginal (std rows),
// flattened to the error vector of estimate_error_norm.
Vec observed_std(const Problem& P, const BNormal& obs) {
  Vec out;
  for (int i = 0; i < P.nblk; ++i) {
    Vec s = P.ops.std_rows(obs.b[i]);
    if (P.bd()) {
      // blockdiag std = vector_norm per row (ssm_impl_blockdiag.py:269-274).
      Real ss = 0;
      for (int j = 0; j < obs.b[i].L.c; ++j) ss += obs.b[i].L(0, j) * obs.b[i].L(0, j);
      s[0] = std::sqrt(ss);
    }
    out.insert(out.end(), s.begin(), s.end());
  }
  return out;
}

// =========================================================== solutions
struct Post {
  BNormal marg;
  BCond cond;  // smoothers only
};
struct Sol {
  Real t = 0;
  BNormal u;
  Post post;
  Vec os;
  long num_steps = 0;
  BCond fx;             // fun_evals (cached linearisation)
  uint32_t lin_set = 0;  // the Monte Carlo handlers' next probe set (the auxiliary key)
  Vec mle_running;      // solver_mle auxiliary
  Real mle_num_data = 0;
};

bool smoother(const Problem& P) { return P.cfg.strategy != STRAT_FILTER; }

// strategy.predict (estimators_and_losses.py:371-373, 526-532, 612-620).
void predict(const Problem& P, const Post& post, const BCond& tr, BNormal& u, Post& pred) {
  if (!smoother(P)) {
    pred.marg = marginalise(P, tr, post.marg);
    u = pred.marg;
    return;
  }
  BCond cond;
  revert(P, tr, post.marg, SOLVE_TRIU, pred.marg, cond);
  if (P.cfg.strategy == STRAT_FIXEDPOINT)
    pred.cond = merge(P, post.cond, cond);
  else
    pred.cond = cond;
  u = pred.marg;
}

// bayes_rule_tree(zeros, rv) (ssm_impl_api.py:112-116) and its variants.
BNormal bayes_rule_zero(const Problem& P, const BCond& fx, const BNormal& rv, BNormal* observed_out) {
  BNormal obs;
  BCond rev;
  revert(P, fx, rv, SOLVE_TRIU, obs, rev);
  BNormal out;
  out.b.resize(P.nblk);
  for (int i = 0; i < P.nblk; ++i) {
    Mat zeros(fx.b[i].noise.m.r, fx.b[i].noise.m.c);
    out.b[i] = P.ops.apply_flat(rev.b[i], zeros);
  }
  if (observed_out) *observed_out = obs;
  return out;
}

Vec prototype_os(const Problem& P) { return Vec(P.nblk, 1.0); }

// solver*.init (solvers.py:515-547, 664-700, 348-393), constraint_init=None.
Sol solver_init(const Problem& P, Real t, const BNormal& init) {
  Sol s;
  s.t = t;
  s.u = init;
  s.post.marg = init;
  if (smoother(P)) s.post.cond = identity_conditional(P, init);
  s.os = prototype_os(P);
  s.num_steps = 0;
  // fun_evals = zeros_like(linearisation) (only its structure matters).
  s.fx = linearize(P, init, t);
  for (auto& c : s.fx.b) {
    for (auto& x : c.A.a) x = 0;
    for (auto& x : c.noise.m.a) x = 0;
    for (auto& x : c.noise.L.a) x = 0;
    for (auto& x : c.tl) x = 0;
    for (auto& x : c.to) x = 0;
  }
  s.mle_running = Vec(P.nblk, 0.0);
  s.mle_num_data = 0.0;
  return s;
}

// Calibration tape (test infrastructure for gradients): in record mode every
// solver_dynamic step appends its output scale, in replay mode the steps use
// the recorded scales instead of recomputing them.  Differentiating the LML
// with a replayed tape is the reference's stop_gradient rule for the
// calibration (solvers.py:568-569): the scales are constants.
enum { TAPE_OFF = 0, TAPE_RECORD = 1, TAPE_REPLAY = 2 };
static thread_local int g_tape_mode = TAPE_OFF;
static thread_local Real* g_tape = nullptr;
static thread_local size_t g_tape_len = 0, g_tape_pos = 0;

// solver_dynamic.step (solvers.py:549-599), solver.step (:702-733),
// solver_mle.step (:395-437).
Sol solver_step(const Problem& P, const Sol& st, Real dt) {
  Sol out;
  const Real tn = st.t + dt;
  if (P.cfg.solver == SOLVER_DYNAMIC) {
    BCond tr = transition(P, dt, prototype_os(P));
    BNormal u = apply_flat(P, tr, st.u);
    BCond fx = linearize(P, u, tn, st.lin_set);
    BNormal observed = marginalise(P, fx, u);
    Vec os = residual_whitened_rms_zero(P, observed);
    if (g_tape_mode != TAPE_OFF) {
      for (auto& x : os) {
        if (g_tape_pos >= g_tape_len) throw std::runtime_error("calibration tape too short");
        if (g_tape_mode == TAPE_RECORD) g_tape[g_tape_pos] = x;
        else x = g_tape[g_tape_pos];
        ++g_tape_pos;
      }
    }
    BCond tr2 = transition(P, dt, os);
    BNormal upred;
    Post pred;
    predict(P, st.post, tr2, upred, pred);
    // solvers.py:576-582: optionally linearise again at the calibrated
    // prediction; the step still reports the initial linearisation (:597)
    const BCond fx_upd = P.cfg.re_linearize ? linearize(P, upred, tn, st.lin_set + 1) : fx;
    BNormal upd = bayes_rule_zero(P, fx_upd, upred, nullptr);
    out.u = upd;
    out.post.marg = upd;
    if (smoother(P)) out.post.cond = pred.cond;
    out.os = os;
    out.fx = fx;
  } else {
    Vec os1 = prototype_os(P);
    BCond tr = transition(P, dt, os1);
    BNormal upred;
    Post pred;
    predict(P, st.post, tr, upred, pred);
    BCond fx = linearize(P, upred, tn, st.lin_set);
    BNormal observed;
    BNormal upd = bayes_rule_zero(P, fx, upred, &observed);
    out.u = upd;
    out.post.marg = upd;
    if (smoother(P)) out.post.cond = pred.cond;
    out.fx = fx;
    if (P.cfg.solver == SOLVER_MLE) {
      Vec new_term = residual_whitened_rms_zero(P, observed);
      out.mle_running.resize(P.nblk);
      for (int i = 0; i < P.nblk; ++i) {
        const Real x1 = std::sqrt(st.mle_num_data / (st.mle_num_data + 1)) * st.mle_running[i];
        const Real x2 = std::sqrt(1 / (st.mle_num_data + 1)) * new_term[i];
        out.mle_running[i] = std::hypot(x1, x2);
      }
      out.mle_num_data = st.mle_num_data + 1;
      out.os = st.os;
    } else {
      out.os = os1;
    }
  }
  out.t = tn;
  out.num_steps = st.num_steps + 1;
  out.lin_set = st.lin_set + lin_sets_per_step(P.cfg);
  return out;
}

// error_residual_std.estimate_error_norm (solvers.py:925-996) with
// error_norm_scale_then_rms / _rms_then_scale (:770-811).
// The std of a state's derivative idx: dense rows idx d .. idx d + d - 1 (per
// row qr_r, ssm_impl_dense.py:141-150), isotropic row idx and blockdiag row
// idx of every block (vector norms, ssm_impl_isotropic.py:190-197 and
// ssm_impl_blockdiag.py:269-274).
Vec state_std_at(const Problem& P, const BNormal& x, int idx) {
  if (P.cfg.fact == FACT_DENSE) {
    const Vec s = P.ops.std_rows(x.b[0]);
    return Vec(s.begin() + idx * P.d, s.begin() + (idx + 1) * P.d);
  }
  Vec out(P.nblk);
  for (int i = 0; i < P.nblk; ++i) {
    Real ss = 0;
    for (int j = 0; j < x.b[i].L.c; ++j) ss += x.b[i].L(idx, j) * x.b[i].L(idx, j);
    out[i] = std::sqrt(ss);
  }
  return out;
}

// error_residual_std.estimate_error_norm (solvers.py:925-996) and
// error_state_std.estimate_error_norm (:1041-1098).  Both extrapolate the
// previous mean with output scale 1 (rv) and take the step's linearisation,
// or re-linearise at rv (re_linearize_before_error).  The residual estimator
// measures the calibrated std of the observed residual (order n = 1); the
// state estimator conditions rv on the residual and measures the calibrated
// std of derivative idx of the result (order n = idx).  error_per_unit_step
// raises n by one.
Real estimate_error_power(const Problem& P, const Sol& prev, const Sol& prop, Real dt) {
  BCond tr = transition(P, dt, prototype_os(P));
  BNormal rv = apply_flat(P, tr, prev.u);
  const BCond fx = P.cfg.error_re_linearize ? linearize(P, rv, prop.t) : prop.fx;
  Vec err, u0, u1;
  int nord;
  if (P.cfg.error_estimator == 1) {
    BNormal observed;
    const BNormal updated = bayes_rule_zero(P, fx, rv, &observed);
    const Vec os = residual_whitened_rms_zero(P, observed);
    const int idx = P.cfg.error_derivative_idx;
    err = state_std_at(P, updated, idx);
    for (size_t i = 0; i < err.size(); ++i) err[i] = os[os.size() == 1 ? 0 : i] * err[i];
    u0 = coeff_of(P, prev.u, idx);
    u1 = coeff_of(P, prop.u, idx);
    nord = idx;
  } else {
    BNormal observed = marginalise(P, fx, rv);
    Vec os = residual_whitened_rms_zero(P, observed);
    rescale_cholesky(P, observed, os);
    err = observed_std(P, observed);
    u0 = u0_of(P, prev.u);
    u1 = u0_of(P, prop.u);
    nord = 1;  // residual_order - 1 for first-order ODEs
  }
  if (P.cfg.error_per_unit_step) nord += 1;
  const int d = P.d;
  Vec ref(d);
  for (int k = 0; k < d; ++k) ref[k] = std::max(std::fabs(u0[k]), std::fabs(u1[k]));
  Vec err_abs(err.size());
  for (size_t i = 0; i < err.size(); ++i) err_abs[i] = err[i] * std::pow(dt, (Real)nord) / factorial_int(nord);
  Real norm;
  if (P.cfg.error_norm == NORM_SCALE_THEN_RMS) {
    Real ss = 0;
    for (int k = 0; k < d; ++k) {
      const Real e = err_abs[err_abs.size() == 1 ? 0 : k];
      const Real r = e / (P.cfg.atol + P.cfg.rtol * std::fabs(ref[k]));
      ss += r * r;
    }
    norm = std::sqrt(ss) / std::sqrt((Real)d);
  } else {
    Real sa = 0, sr = 0;
    for (Real e : err_abs) sa += e * e;
    for (Real r : ref) sr += r * r;
    const Real na = std::sqrt(sa) / std::sqrt((Real)err_abs.size());
    const Real nr = std::sqrt(sr) / std::sqrt((Real)ref.size());
    norm = na / (P.cfg.atol + P.cfg.rtol * nr);
  }
  return std::pow(norm, -1.0 / (Real)P.N);
}

// controllers.py:66-84 control_integral / :24-63 proportional-integral.
Real control_apply(const Problem& P, Real dt, Real& cstate, Real error_power) {
  const auto& c = P.cfg;
  if (c.control == CONTROL_INTEGRAL) {
    const Real ratio = c.safety * error_power;
    const Real clipped = std::min(ratio, c.factor_max);
    return std::max(c.factor_min, clipped) * dt;
  }
  const Real gi = std::pow(error_power, c.exponent_integral);
  const Real gp = std::pow(error_power / cstate, c.exponent_proportional);
  const Real ratio = c.safety * gi * gp;
  const Real clipped = std::min(ratio, c.factor_max);
  const Real sf = std::max(c.factor_min, clipped);
  if (error_power >= 1.0) cstate = error_power;
  return sf * dt;
}

// ProbabilisticSolver.interpolate_fwd / _at_t1 (solvers.py:205-315) with
// the strategies' interpolate_fwd (estimators_and_losses.py:393-421,
// 534-591, 622-672, 711-717).
void interpolate_fwd(const Problem& P, Real t, const Sol& from, const Sol& to, Sol& interpolated, Sol& step_from,
                     Sol& interp_from) {
  BCond tr0 = transition(P, t - from.t, to.os);
  BCond tr1 = transition(P, to.t - t, to.os);
  Post interp_post, step_post, from_post;
  BNormal marg;
  if (P.cfg.strategy == STRAT_FILTER) {
    BNormal u;
    predict(P, from.post, tr0, u, interp_post);
    marg = interp_post.marg;
    step_post = to.post;
    from_post = interp_post;
  } else if (P.cfg.strategy == STRAT_FIXEDPOINT) {
    BNormal u;
    Post ext_t;
    predict(P, from.post, tr0, u, ext_t);
    Post previous_new;
    previous_new.marg = ext_t.marg;
    previous_new.cond = identity_conditional(P, from.post.marg);
    Post ext_t1;
    predict(P, previous_new, tr1, u, ext_t1);
    interp_post = ext_t;
    step_post.marg = to.post.marg;
    step_post.cond = ext_t1.cond;
    from_post = previous_new;
    marg = ext_t.marg;
  } else {
    BNormal u;
    Post sol_t, ext_t1;
    predict(P, from.post, tr0, u, sol_t);
    predict(P, sol_t, tr1, u, ext_t1);
    step_post.marg = to.post.marg;
    step_post.cond = ext_t1.cond;
    from_post = sol_t;
    interp_post = sol_t;
    marg = sol_t.marg;
  }
  step_from = to;
  step_from.post = step_post;
  interpolated = to;
  interpolated.t = t;
  interpolated.post = interp_post;
  interpolated.u = marg;
  interp_from = from;
  interp_from.t = t;
  interp_from.post = from_post;
}

void interpolate_fwd_at_t1(const Problem& P, const Sol& from, const Sol& to, Sol& sol, Sol& step_from, Sol& interp_from) {
  Post interpolated = to.post, step_post = to.post, from_post = to.post;
  BNormal marg = to.post.marg;
  if (P.cfg.strategy == STRAT_FIXEDPOINT) {
    Post resume;
    resume.marg = to.post.marg;
    resume.cond = identity_conditional(P, to.post.marg);
    step_post = resume;
    from_post = resume;
  }
  interp_from = from;
  interp_from.t = to.t;
  interp_from.post = from_post;
  sol = to;
  sol.post = interpolated;
  sol.u = marg;
  step_from = to;
  step_from.post = step_post;
}

// ---------------------------------------------------------- outputs
struct Outputs {
  std::vector<Sol> sols;  // the per-checkpoint (or per-step) solutions
  Sol sol0, sol1;
  long attempts = 0, accepted = 0;
  int status = 0;
};

enum { ST_OK = 0, ST_MAX_ATTEMPTS = 1, ST_NONFINITE = 2, ST_DT_UNDERFLOW = 3 };
// Bound on the attempts of one solve (the kernels use the same bound, so that
// every persistent wave reaches its exit in bounded time).
constexpr long kMaxTotalAttempts = 1L << 22;

// solve_adaptive_save_at (solvers_via_adaptive_steps.py:46-148) with the
// RejectionLoop (:196-375).
void solve_save_at(const Problem& P, const BNormal& init, const Real* save_at, int K, Outputs& out) {
  const auto& c = P.cfg;
  Sol sol0 = solver_init(P, save_at[0], init);
  out.sol0 = sol0;
  Real dt = c.dt0, cstate = 1.0;
  Sol step_from = sol0, interp_from = sol0;
  for (int ck = 1; ck < K && out.status == ST_OK; ++ck) {
    const Real t1 = save_at[ck];
    bool do_continue = true;
    Sol solution;
    while (do_continue) {
      if (step_from.t + c.eps < t1) {
        // RejectionLoop.step: attempt until error_power >= 1.
        Real ep = 0.9;
        Sol proposed;
        int attempts = 0;
        while (ep < 1.0) {
          Real dtc = dt;
          if (c.clip_dt) dtc = std::min(dtc, t1 - step_from.t);
          proposed = solver_step(P, step_from, dtc);
          ep = estimate_error_power(P, step_from, proposed, dtc);
          dt = control_apply(P, dtc, cstate, ep);
          ++attempts;
          ++out.attempts;
          if (!std::isfinite(ep) || !std::isfinite(dt)) {
            out.status = ST_NONFINITE;
            break;
          }
          // Failure semantics the reference lacks (it would loop forever):
          // a rejected attempt whose next step cannot advance t, or an
          // accepted one that did not, ends the solve with DT_UNDERFLOW;
          // more than kMaxTotalAttempts attempts end it with MAX_ATTEMPTS.
          if (ep < 1.0 && (dt == 0.0 || step_from.t + dt == step_from.t)) {
            out.status = ST_DT_UNDERFLOW;
            break;
          }
          if (ep >= 1.0 && proposed.t == step_from.t) {
            out.status = ST_DT_UNDERFLOW;
            break;
          }
          if ((c.max_attempts > 0 && attempts >= c.max_attempts && ep < 1.0) || out.attempts >= kMaxTotalAttempts) {
            out.status = ST_MAX_ATTEMPTS;
            break;
          }
        }
        if (out.status != ST_OK) break;
        ++out.accepted;
        interp_from = step_from;
        step_from = proposed;
      }
      const bool before = step_from.t + c.eps < t1;
      const bool after = step_from.t > t1 + c.eps;
      if (before) {
        solution = step_from;
      } else if (after) {
        Sol sf, ifr;
        interpolate_fwd(P, t1, interp_from, step_from, solution, sf, ifr);
        step_from = sf;
        interp_from = ifr;
      } else {
        Sol sf, ifr;
        interpolate_fwd_at_t1(P, interp_from, step_from, solution, sf, ifr);
        step_from = sf;
        interp_from = ifr;
      }
      do_continue = step_from.t + c.eps < t1;
    }
    out.sols.push_back(solution);
  }
  out.sol1 = step_from;
}

// solve_fixed_grid (solvers_via_fixed_steps.py:11-34).
void solve_fixed_grid(const Problem& P, const BNormal& init, const Real* grid, int K, Outputs& out) {
  Sol s = solver_init(P, grid[0], init);
  out.sol0 = s;
  for (int k = 1; k < K; ++k) {
    const Real dt = grid[k] - grid[k - 1];
    s = solver_step(P, s, dt);
    ++out.attempts;
    ++out.accepted;
    out.sols.push_back(s);
  }
  out.sol1 = s;
}

// util/test_util.py:10-66 solve_adaptive_save_every_step: call the
// rejection loop until step_from.t >= t1 and keep every returned solution.
void solve_save_every_step(const Problem& P, const BNormal& init, Real t0, Real t1, Outputs& out) {
  const auto& c = P.cfg;
  Sol sol0 = solver_init(P, t0, init);
  out.sol0 = sol0;
  Real dt = c.dt0, cstate = 1.0;
  Sol step_from = sol0, interp_from = sol0;
  while (step_from.t < t1 && out.status == ST_OK) {
    Sol solution;
    if (step_from.t + c.eps < t1) {
      Real ep = 0.9;
      Sol proposed;
      while (ep < 1.0) {
        Real dtc = dt;
        if (c.clip_dt) dtc = std::min(dtc, t1 - step_from.t);
        proposed = solver_step(P, step_from, dtc);
        ep = estimate_error_power(P, step_from, proposed, dtc);
        dt = control_apply(P, dtc, cstate, ep);
        ++out.attempts;
        if (!std::isfinite(ep)) {
          out.status = ST_NONFINITE;
          break;
        }
        if ((ep < 1.0 && (dt == 0.0 || step_from.t + dt == step_from.t)) || (ep >= 1.0 && proposed.t == step_from.t)) {
          out.status = ST_DT_UNDERFLOW;
          break;
        }
        if (out.attempts >= kMaxTotalAttempts) {
          out.status = ST_MAX_ATTEMPTS;
          break;
        }
      }
      if (out.status != ST_OK) break;
      ++out.accepted;
      interp_from = step_from;
      step_from = proposed;
    }
    const bool before = step_from.t + c.eps < t1;
    const bool after = step_from.t > t1 + c.eps;
    if (before) {
      solution = step_from;
    } else if (after) {
      Sol sf, ifr;
      interpolate_fwd(P, t1, interp_from, step_from, solution, sf, ifr);
      step_from = sf;
      interp_from = ifr;
    } else {
      Sol sf, ifr;
      interpolate_fwd_at_t1(P, interp_from, step_from, solution, sf, ifr);
      step_from = sf;
      interp_from = ifr;
    }
    out.sols.push_back(solution);
  }
  out.sol1 = step_from;
}

// Forward pass of the LML configuration: the fixed-point smoother on a fixed
// grid with checkpoints at every `every`-th grid point (interp_fwd_at_t1
// semantics at each checkpoint).
void solve_fixed_grid_checkpoints(const Problem& P, const BNormal& init, const Real* grid, int K, int every,
                                  Outputs& out) {
  Sol s = solver_init(P, grid[0], init);
  out.sol0 = s;
  Sol interp_from = s;
  for (int k = 1; k < K; ++k) {
    const Real dt = grid[k] - grid[k - 1];
    Sol prop = solver_step(P, s, dt);
    ++out.attempts;
    ++out.accepted;
    Sol prev = s;
    s = prop;
    if (k % every == 0) {
      Sol sol, sf, ifr;
      interpolate_fwd_at_t1(P, prev, s, sol, sf, ifr);
      out.sols.push_back(sol);
      s = sf;
    }
  }
  out.sol1 = s;
}

// userfriendly_output + strategy.finalize (solvers.py:439-480, 601-633,
// 735-767; estimators_and_losses.py:380-391, 437-470).
struct Final {
  std::vector<Real> t;
  std::vector<BNormal> u;          // K marginals
  std::vector<BNormal> filtering;  // smoothers: K filtering marginals
  std::vector<BCond> conds;        // smoothers: K-1 backward conditionals
  BNormal terminal;                // smoothers: full_posterior.marginal
  std::vector<Vec> os;             // K output scales
  std::vector<long> num_steps;     // K-1
};

void finalize(const Problem& P, const Outputs& o, Final& F) {
  const int Km1 = (int)o.sols.size();
  Vec scale = prototype_os(P);
  if (P.cfg.solver == SOLVER_MLE) {
    scale = o.sol1.mle_running;
    if (P.cfg.correct_asymptotic_underconfidence)
      for (auto& x : scale) x = x / std::sqrt((Real)o.sols.back().num_steps);
  }
  F.t.push_back(o.sol0.t);
  for (auto& s : o.sols) F.t.push_back(s.t);
  if (P.cfg.solver == SOLVER_DYNAMIC) {
    F.os.push_back(o.sol0.os);
    for (auto& s : o.sols) F.os.push_back(s.os);
  } else {
    // solvers.py:755 / :469: ones_like(solution.output_scale) * scale has
    // one entry per checkpoint after t0 (K-1 rows, unlike solver_dynamic).
    for (int i = 0; i < Km1; ++i) F.os.push_back(scale);
  }
  for (auto& s : o.sols) F.num_steps.push_back(s.num_steps);
  if (!smoother(P)) {
    BNormal p0 = o.sol0.post.marg;
    rescale_cholesky(P, p0, scale);
    F.u.push_back(p0);
    for (auto& s : o.sols) {
      BNormal p = s.post.marg;
      rescale_cholesky(P, p, scale);
      F.u.push_back(p);
    }
    return;
  }
  Post p0 = o.sol0.post, p1 = o.sol1.post;
  rescale_cholesky(P, p0.marg, scale);
  rescale_noise(P, p0.cond, scale);
  rescale_cholesky(P, p1.marg, scale);
  rescale_noise(P, p1.cond, scale);
  std::vector<Post> post(Km1);
  for (int i = 0; i < Km1; ++i) {
    post[i] = o.sols[i].post;
    rescale_cholesky(P, post[i].marg, scale);
    rescale_noise(P, post[i].cond, scale);
  }
  BNormal rv_at_t1 = marginalise(P, p1.cond, p1.marg);
  F.terminal = rv_at_t1;
  // MarkovSequence.evaluate_marginals (estimators_and_losses.py:156-178).
  std::vector<BNormal> margs(Km1);
  BNormal x = rv_at_t1;
  for (int i = Km1 - 1; i >= 0; --i) {
    x = marginalise(P, post[i].cond, x);
    margs[i] = x;
  }
  F.u = margs;
  F.u.push_back(rv_at_t1);
  F.filtering.push_back(p0.marg);
  for (int i = 0; i < Km1; ++i) {
    F.filtering.push_back(post[i].marg);
    F.conds.push_back(post[i].cond);
  }
}

// The observation model marginal.to_derivative(i, std) (ssm_impl_dense.py:
// 206-217, ssm_impl_isotropic.py:271-278, ssm_impl_blockdiag.py:367-377) with
// i = cfg.tcoeff_index: dense selects rows i d + k of the derivative-major
// state, isotropic and blockdiag row i.
BCond observation_model(const Problem& P, const Real* std_j, int nstd) {
  BCond out;
  out.b.resize(P.nblk);
  const int d = P.d, N = P.N, ix = P.cfg.tcoeff_index;
  if (P.cfg.fact == FACT_DENSE) {
    Cond& C = out.b[0];
    C.A = Mat(d, P.n);
    for (int k = 0; k < d; ++k) C.A(k, ix * d + k) = 1.0;
    C.noise.m = Mat(d, 1);
    C.noise.L = Mat(d, d);
    for (int k = 0; k < d; ++k) C.noise.L(k, k) = std_j[nstd == 1 ? 0 : k];
    C.tl = ones(P.n);
    C.to = ones(d);
  } else if (P.cfg.fact == FACT_ISO) {
    Cond& C = out.b[0];
    C.A = Mat(1, N);
    C.A(0, ix) = 1.0;
    C.noise.m = Mat(1, d);
    C.noise.L = Mat(1, 1);
    C.noise.L(0, 0) = std_j[0];
    C.tl = ones(N);
    C.to = ones(1);
  } else {
    for (int k = 0; k < d; ++k) {
      Cond& C = out.b[k];
      C.A = Mat(1, N);
      C.A(0, ix) = 1.0;
      C.noise.m = Mat(1, 1);
      C.noise.L = Mat(1, 1);
      C.noise.L(0, 0) = std_j[nstd == 1 ? 0 : k];
      C.tl = ones(N);
      C.to = ones(1);
    }
  }
  return out;
}

// bayes_rule_and_logpdf_tree (ssm_impl_api.py:118-124) with lstsq gains.
Real bayes_logpdf(const Problem& P, const BCond& model, const Real* y, const BNormal& rv, BNormal& updated) {
  BNormal obs;
  BCond rev;
  revert(P, model, rv, P.cfg.lml_solve == 1 ? SOLVE_TRIU : SOLVE_LSTSQ, obs, rev);
  Real lp = 0.0;
  updated.b.resize(P.nblk);
  for (int i = 0; i < P.nblk; ++i) {
    Mat data(obs.b[i].m.r, obs.b[i].m.c);
    if (P.cfg.fact == FACT_BLOCKDIAG)
      data(0, 0) = y[i];
    else
      for (size_t e = 0; e < data.a.size(); ++e) data.a[e] = y[e];
    if (P.cfg.fact == FACT_BLOCKDIAG) {
      // logpdf_scalar_flat per block (ssm_impl_blockdiag.py:302-316).
      SSMBase iso;
      iso.fact = FACT_ISO;
      lp += iso.logpdf(obs.b[i], data);
    } else {
      lp += P.ops.logpdf(obs.b[i], data);
    }
    updated.b[i] = P.ops.apply_flat(rev.b[i], data);
  }
  return lp;
}

// MarkovSequence.evaluate_lml (estimators_and_losses.py:180-218) via
// loss_lml_timeseries (:53-105): data y (K x d), std (K x nstd).
Real evaluate_lml(const Problem& P, const Final& F, const Real* y, const Real* stds, int nstd, int average) {
  const int K = (int)F.conds.size() + 1, d = P.d;
  BNormal upd;
  BCond m0 = observation_model(P, stds + (size_t)(K - 1) * nstd, nstd);
  Real logpdf = bayes_logpdf(P, m0, y + (size_t)(K - 1) * d, F.terminal, upd);
  Real num = 1.0;
  BNormal rv = upd;
  for (int j = K - 2; j >= 0; --j) {
    BNormal pred = marginalise(P, F.conds[j], rv);
    BCond mj = observation_model(P, stds + (size_t)j * nstd, nstd);
    BNormal corrected;
    Real lpn = bayes_logpdf(P, mj, y + (size_t)j * d, pred, corrected);
    if (average)
      logpdf = (logpdf * num + lpn) / (num + 1);
    else
      logpdf = logpdf + lpn;
    num += 1;
    rv = corrected;
  }
  return logpdf;
}

// Prior init normal from Taylor coefficients (tcoeffs K x d, derivatives)
// and is_exact (std zero) or inexact std.
BNormal prior_init(const Problem& P, const Real* tc, Real init_std) {
  BNormal out;
  out.b.resize(P.nblk);
  const int N = P.N, d = P.d;
  if (P.cfg.fact == FACT_DENSE) {
    Normal& x = out.b[0];
    x.m = Mat(P.n, 1);
    x.L = Mat(P.n, P.n);
    for (int i = 0; i < N; ++i)
      for (int k = 0; k < d; ++k) x.m(i * d + k, 0) = tc[i * d + k];
    for (int i = 0; i < P.n; ++i) x.L(i, i) = P.cfg.init_std_rows ? P.cfg.init_std_rows[i] : init_std;
  } else if (P.cfg.fact == FACT_ISO) {
    Normal& x = out.b[0];
    x.m = Mat(N, d);
    x.L = Mat(N, N);
    for (int i = 0; i < N; ++i)
      for (int k = 0; k < d; ++k) x.m(i, k) = tc[i * d + k];
    for (int i = 0; i < N; ++i) x.L(i, i) = P.cfg.init_std_rows ? P.cfg.init_std_rows[i] : init_std;
  } else {
    for (int k = 0; k < d; ++k) {
      Normal& x = out.b[k];
      x.m = Mat(N, 1);
      x.L = Mat(N, N);
      for (int i = 0; i < N; ++i) x.m(i, 0) = tc[i * d + k];
      for (int i = 0; i < N; ++i) x.L(i, i) = init_std;
    }
  }
  return out;
}

// Flat layouts of a (blocked) normal: dense mean n, chol n*n; isotropic
// mean N*d, chol N*N; blockdiag mean d*N, chol d*N*N.
int mean_size(const Problem& P) { return P.cfg.fact == FACT_BLOCKDIAG ? P.d * P.N : P.n * (P.cfg.fact == FACT_ISO ? P.d : 1); }
int chol_size(const Problem& P) { return P.nblk * P.n * P.n; }

void write_normal(const Problem& P, const BNormal& x, Real* mean, Real* chol) {
  if (P.cfg.fact == FACT_BLOCKDIAG) {
    for (int k = 0; k < P.d; ++k)
      for (int i = 0; i < P.N; ++i) {
        if (mean) mean[k * P.N + i] = x.b[k].m(i, 0);
        if (chol)
          for (int j = 0; j < P.N; ++j) chol[(size_t)k * P.N * P.N + i * P.N + j] = x.b[k].L(i, j);
      }
    return;
  }
  if (mean) std::memcpy(mean, x.b[0].m.a.data(), sizeof(Real) * x.b[0].m.a.size());
  if (chol) std::memcpy(chol, x.b[0].L.a.data(), sizeof(Real) * x.b[0].L.a.size());
}

BNormal read_normal(const Problem& P, const Real* mean, const Real* chol) {
  BNormal x;
  x.b.resize(P.nblk);
  if (P.cfg.fact == FACT_BLOCKDIAG) {
    for (int k = 0; k < P.d; ++k) {
      x.b[k].m = Mat(P.N, 1);
      x.b[k].L = Mat(P.N, P.N);
      for (int i = 0; i < P.N; ++i) {
        x.b[k].m(i, 0) = mean[k * P.N + i];
        for (int j = 0; j < P.N; ++j) x.b[k].L(i, j) = chol[(size_t)k * P.N * P.N + i * P.N + j];
      }
    }
    return x;
  }
  const int cols = P.cfg.fact == FACT_ISO ? P.d : 1;
  x.b[0].m = Mat(P.n, cols);
  x.b[0].L = Mat(P.n, P.n);
  std::memcpy(x.b[0].m.a.data(), mean, sizeof(Real) * x.b[0].m.a.size());
  std::memcpy(x.b[0].L.a.data(), chol, sizeof(Real) * x.b[0].L.a.size());
  return x;
}

// A conditional's flat layout: A (nblk*n*n), noise mean (mean_size), noise
// chol (chol_size), to_latent (nblk*n), to_observed (nblk*n).
void write_cond(const Problem& P, const BCond& c, Real* A, Real* nm, Real* nc, Real* tl, Real* to) {
  BNormal noise;
  noise.b.resize(P.nblk);
  for (int i = 0; i < P.nblk; ++i) noise.b[i] = c.b[i].noise;
  write_normal(P, noise, nm, nc);
  for (int i = 0; i < P.nblk; ++i) {
    std::memcpy(A + (size_t)i * P.n * P.n, c.b[i].A.a.data(), sizeof(Real) * P.n * P.n);
    std::memcpy(tl + (size_t)i * P.n, c.b[i].tl.data(), sizeof(Real) * P.n);
    std::memcpy(to + (size_t)i * P.n, c.b[i].to.data(), sizeof(Real) * P.n);
  }
}

BCond read_cond(const Problem& P, const Real* A, const Real* nm, const Real* nc, const Real* tl, const Real* to) {
  BCond c;
  c.b.resize(P.nblk);
  const BNormal noise = read_normal(P, nm, nc);
  for (int i = 0; i < P.nblk; ++i) {
    c.b[i].noise = noise.b[i];
    c.b[i].A = Mat(P.n, P.n);
    std::memcpy(c.b[i].A.a.data(), A + (size_t)i * P.n * P.n, sizeof(Real) * P.n * P.n);
    c.b[i].tl = Vec(tl + (size_t)i * P.n, tl + (size_t)(i + 1) * P.n);
    c.b[i].to = Vec(to + (size_t)i * P.n, to + (size_t)(i + 1) * P.n);
  }
  return c;
}

// Normal.sample_flat with given standard normal draws z: dense
// (ssm_impl_dense.py:192-194) m + L z with z of length n; isotropic
// (ssm_impl_isotropic.py:255-258) m + (L z)[:, None] with z of length N, the
// same draw for every one of the d columns.
Normal sample_with(const Normal& x, const Real* z) {
  Normal out = x;
  for (int i = 0; i < x.L.r; ++i) {
    Real acc = 0.0;
    for (int j = 0; j < x.L.c; ++j) acc += x.L(i, j) * z[j];
    for (int k = 0; k < x.m.c; ++k) out.m(i, k) = x.m(i, k) + acc;
  }
  return out;
}

}  // namespace

// =============================================================== C API
extern "C" {

struct pdqo_out {
  // Caller-allocated; any pointer may be null.
  Real* t;            // K
  Real* mean;         // K * mean_size
  Real* chol;         // K * chol_size
  Real* os;           // K * nblk
  long* num_steps;      // K-1
  Real* filt_mean;    // smoothers: K * mean_size
  Real* filt_chol;    // K * chol_size
  Real* cond_A;       // smoothers: (K-1) * nblk*n*n
  Real* cond_mean;    // (K-1) * mean_size
  Real* cond_chol;    // (K-1) * chol_size
  Real* cond_tl;      // (K-1) * nblk*n
  Real* cond_to;      // (K-1) * nblk*n
  Real* term_mean;    // smoothers: mean_size
  Real* term_chol;    // chol_size
  long attempts, accepted;
  int status;
};

int pdqo_version(void) { return 1; }

// Size of the scalar type every array and the configuration use.
int pdqo_real_size(void) { return (int)sizeof(Real); }

void pdqo_system_matrices(int nu, Real* A, Real* Q) {
  Mat A1, Q1;
  system_matrices_1d_iwp(nu, A1, Q1);
  std::memcpy(A, A1.a.data(), sizeof(Real) * A1.a.size());
  std::memcpy(Q, Q1.a.data(), sizeof(Real) * Q1.a.size());
}

void pdqo_qr_r(int m, int n, const Real* M, Real* R) {
  Mat A(m, n);
  std::memcpy(A.a.data(), M, sizeof(Real) * m * n);
  Mat r = qr_r(A);
  std::memcpy(R, r.a.data(), sizeof(Real) * r.a.size());
}

// lstsq_svd (backend/linalg.py:60-61) of A (m x n) and B (m x nb): X (n x nb).
void pdqo_lstsq_svd(int m, int n, const Real* A, int nb, const Real* B, Real* X) {
  Mat a(m, n), b(m, nb);
  std::memcpy(a.a.data(), A, sizeof(Real) * m * n);
  std::memcpy(b.a.data(), B, sizeof(Real) * m * nb);
  Mat x = lstsq_svd(a, b);
  std::memcpy(X, x.a.data(), sizeof(Real) * n * nb);
}

void pdqo_revert_conditional(int n, int k, const Real* R_X_F, const Real* R_X, const Real* R_YX, int mode,
                             Real* R_Y, Real* R_XY, Real* G) {
  Mat a(n, k), b(n, n), c(k, k);
  std::memcpy(a.a.data(), R_X_F, sizeof(Real) * n * k);
  std::memcpy(b.a.data(), R_X, sizeof(Real) * n * n);
  std::memcpy(c.a.data(), R_YX, sizeof(Real) * k * k);
  Reverted r = revert_conditional(a, b, c, mode);
  std::memcpy(R_Y, r.R_Y.a.data(), sizeof(Real) * k * k);
  std::memcpy(R_XY, r.R_XY.a.data(), sizeof(Real) * n * n);
  std::memcpy(G, r.G.a.data(), sizeof(Real) * n * k);
}

// IWP transition conditional for one block (dense: full n).
void pdqo_transition(const pdqo_cfg* cfg, Real dt, const Real* os, Real* A, Real* noise_chol, Real* tl,
                     Real* to) {
  Problem P;
  P.setup(*cfg, nullptr);
  Vec osv(P.nblk);
  for (int i = 0; i < P.nblk; ++i) osv[i] = os[i];
  BCond c = transition(P, dt, osv);
  std::memcpy(A, c.b[0].A.a.data(), sizeof(Real) * P.n * P.n);
  std::memcpy(noise_chol, c.b[0].noise.L.a.data(), sizeof(Real) * P.n * P.n);
  std::memcpy(tl, c.b[0].tl.data(), sizeof(Real) * P.n);
  std::memcpy(to, c.b[0].to.data(), sizeof(Real) * P.n);
}

// gram_util.exp_gram_cholesky(pade_and_legendre_9) of A (n x n) and B (n x m).
void pdqo_exp_gram(int n, int m, const Real* A, const Real* B, Real* eA, Real* U) {
  Mat a(n, n), b(n, m), e, u;
  std::memcpy(a.a.data(), A, sizeof(Real) * n * n);
  std::memcpy(b.a.data(), B, sizeof(Real) * n * m);
  exp_gram(a, b, e, u);
  std::memcpy(eA, e.a.data(), sizeof(Real) * n * n);
  std::memcpy(U, u.a.data(), sizeof(Real) * n * n);
}

Real pdqo_logpdf(int fact, int rows, int cols, const Real* mean, const Real* chol, const Real* u) {
  SSMBase ops;
  ops.fact = fact == FACT_ISO ? FACT_ISO : FACT_DENSE;
  Normal x;
  x.m = Mat(rows, cols);
  x.L = Mat(rows, rows);
  std::memcpy(x.m.a.data(), mean, sizeof(Real) * rows * cols);
  std::memcpy(x.L.a.data(), chol, sizeof(Real) * rows * rows);
  Mat uu(rows, cols);
  std::memcpy(uu.a.data(), u, sizeof(Real) * rows * cols);
  return ops.logpdf(x, uu);
}

Real pdqo_control(const pdqo_cfg* cfg, Real dt, Real* state, Real error_power) {
  Problem P;
  P.cfg = *cfg;
  return control_apply(P, dt, *state, error_power);
}

void pdqo_taylor_coefficients(int vf_id, int d, const Real* params, const Real* y0, Real t0, int num,
                              Real* out) {
  Vf vf;
  vf.id = vf_id;
  vf.d = d;
  vf.prm = params;
  taylor_coefficients(vf, y0, t0, num, out);
}

void pdqo_vf(int vf_id, int d, const Real* params, const Real* y, Real t, Real* f, Real* J) {
  Vf vf;
  vf.id = vf_id;
  vf.d = d;
  vf.prm = params;
  if (f) vf.eval(y, t, f);
  if (J) vf.jac(y, t, J);
}

// The Monte Carlo handlers' pieces: one probe entry, and the diagonal
// estimates of the vector field's Jacobian at y (out: d entries).
Real pdqo_rademacher(uint32_t seed, uint32_t set, uint32_t probe, uint32_t coord) {
  return rademacher(seed, set, probe, coord);
}

void pdqo_mc_diag(int vf_id, int d, const Real* params, const Real* y, Real t, uint32_t seed, uint32_t set,
                  int probes, int fwd, Real* out) {
  Vf vf;
  vf.id = vf_id;
  vf.d = d;
  vf.prm = params;
  std::vector<Real> J((size_t)d * d);
  vf.jac(y, t, J.data());
  for (int k = 0; k < d; ++k) out[k] = mc_diag(J.data(), d, k, seed, set, probes, fwd != 0);
}

static void write_outputs(const Problem& P, const Outputs& o, const Final& F, pdqo_out* out) {
  const int K = (int)F.t.size();
  const int ms = mean_size(P), cs = chol_size(P), nn = P.nblk * P.n * P.n, nv = P.nblk * P.n;
  for (int i = 0; i < K; ++i) {
    if (out->t) out->t[i] = F.t[i];
    write_normal(P, F.u[i], out->mean ? out->mean + (size_t)i * ms : nullptr,
                 out->chol ? out->chol + (size_t)i * cs : nullptr);
    if (out->os && i < (int)F.os.size())
      for (int b = 0; b < P.nblk; ++b) out->os[(size_t)i * P.nblk + b] = F.os[i][b];
    if (i < K - 1 && out->num_steps) out->num_steps[i] = F.num_steps[i];
  }
  if (smoother(P)) {
    for (int i = 0; i < K; ++i)
      write_normal(P, F.filtering[i], out->filt_mean ? out->filt_mean + (size_t)i * ms : nullptr,
                   out->filt_chol ? out->filt_chol + (size_t)i * cs : nullptr);
    if (out->cond_A)
      for (int i = 0; i < K - 1; ++i)
        write_cond(P, F.conds[i], out->cond_A + (size_t)i * nn, out->cond_mean + (size_t)i * ms,
                   out->cond_chol + (size_t)i * cs, out->cond_tl + (size_t)i * nv, out->cond_to + (size_t)i * nv);
    write_normal(P, F.terminal, out->term_mean, out->term_chol);
  }
  out->attempts = o.attempts;
  out->accepted = o.accepted;
  out->status = o.status;
}

// One adaptive save_at solve.  tcoeffs: (nu+1) x d derivatives.
int pdqo_solve_save_at(const pdqo_cfg* cfg, const Real* params, const Real* tcoeffs, Real init_std,
                       const Real* save_at, int K, pdqo_out* out) {
  Problem P;
  P.setup(*cfg, params);
  BNormal init = prior_init(P, tcoeffs, init_std);
  Outputs o;
  solve_save_at(P, init, save_at, K, o);
  if (o.status != ST_OK) {
    out->attempts = o.attempts;
    out->accepted = o.accepted;
    out->status = o.status;
    return o.status;
  }
  Final F;
  finalize(P, o, F);
  write_outputs(P, o, F, out);
  return 0;
}

int pdqo_solve_fixed_grid(const pdqo_cfg* cfg, const Real* params, const Real* tcoeffs, Real init_std,
                          const Real* grid, int K, pdqo_out* out) {
  Problem P;
  P.setup(*cfg, params);
  BNormal init = prior_init(P, tcoeffs, init_std);
  Outputs o;
  solve_fixed_grid(P, init, grid, K, o);
  Final F;
  finalize(P, o, F);
  write_outputs(P, o, F, out);
  return 0;
}

// Marginal log-likelihood of data (Kobs x d) at every `every`-th grid point
// (grid[0] included), fixed-point smoother on the fixed grid.
Real pdqo_lml(const pdqo_cfg* cfg, const Real* params, const Real* tcoeffs, Real init_std, const Real* grid,
                int K, int every, const Real* data, const Real* stds, int nstd, int average) {
  Problem P;
  P.setup(*cfg, params);
  BNormal init = prior_init(P, tcoeffs, init_std);
  Outputs o;
  solve_fixed_grid_checkpoints(P, init, grid, K, every, o);
  Final F;
  finalize(P, o, F);
  return evaluate_lml(P, F, data, stds, nstd, average);
}

// LML as a function of (vector-field params, y0): Taylor coefficients are
// recomputed from y0 (jet init, num = nu).
Real pdqo_lml_from_y0(const pdqo_cfg* cfg, const Real* params, const Real* y0, const Real* grid, int K,
                        int every, const Real* data, const Real* stds, int nstd, int average) {
  std::vector<Real> tc((size_t)(cfg->nu + 1) * cfg->d);
  pdqo_taylor_coefficients(cfg->vf_id, cfg->d, params, y0, grid[0], cfg->nu, tc.data());
  return pdqo_lml(cfg, params, tc.data(), 0.0, grid, K, every, data, stds, nstd, average);
}

// ProbabilisticSolver.offgrid_marginals (solvers.py:149-203) for one query
// time t in (t0, t1) of a solution grid, given the output scale of the
// right-hand checkpoint (`os`, nblk entries):
// * filter (estimators_and_losses.py:403-414): predict the marginal at t0
//   (m0, L0) to t;
// * fixed-interval smoother (:674-707): revert the filtering marginal at t0
//   (m0, L0) to t and on to t1, then marginalise the smoothed marginal at t1
//   (m1, L1) through the backward conditional t1 -> t.
// The fixed-point smoother is not suitable (the reference raises).  Returns
// 0, or 2 for an unsuitable strategy.
int pdqo_offgrid_marginal(const pdqo_cfg* cfg, const Real* params, Real t0, Real t1, Real t, const Real* os,
                          const Real* m0, const Real* L0, const Real* m1, const Real* L1, Real* m_out, Real* L_out) {
  Problem P;
  P.setup(*cfg, params);
  if (P.cfg.strategy == STRAT_FIXEDPOINT) return 2;
  const Vec osv(os, os + P.nblk);
  const BNormal x0 = read_normal(P, m0, L0);
  const BCond tr0 = transition(P, t - t0, osv);
  BNormal out;
  if (P.cfg.strategy == STRAT_FILTER) {
    out = marginalise(P, tr0, x0);
  } else {
    BNormal xt, x1p;
    BCond c_t_t0, c_t1_t;
    revert(P, tr0, x0, SOLVE_TRIU, xt, c_t_t0);
    const BCond tr1 = transition(P, t1 - t, osv);
    revert(P, tr1, xt, SOLVE_TRIU, x1p, c_t1_t);
    out = marginalise(P, c_t1_t, read_normal(P, m1, L1));
  }
  write_normal(P, out, m_out, L_out);
  return 0;
}

// loss_lml_terminal_values (estimators_and_losses.py:20-50): the log-density
// of the data u under marginals.to_derivative(idx, std).marginalise(
// marginals), i.e. derivative idx of the marginal observed with noise std
// (std: 1 entry, or one per block / coordinate as the factorisation's
// observation model takes them).
Real pdqo_lml_terminal(const pdqo_cfg* cfg, const Real* params, const Real* mean, const Real* chol, const Real* u,
                       const Real* std, int nstd, int idx) {
  Problem P;
  P.setup(*cfg, params);
  const BNormal x = read_normal(P, mean, chol);
  BCond model = observation_model(P, std, nstd);  // selects derivative 0
  for (auto& C : model.b) {
    if (P.cfg.fact == FACT_DENSE) {
      C.A = Mat(P.d, P.n);
      for (int k = 0; k < P.d; ++k) C.A(k, idx * P.d + k) = 1.0;
    } else {
      C.A = Mat(1, P.N);
      C.A(0, idx) = 1.0;
    }
  }
  const BNormal obs = marginalise(P, model, x);
  Real lp = 0.0;
  for (int i = 0; i < P.nblk; ++i) {
    Mat data(obs.b[i].m.r, obs.b[i].m.c);
    if (P.cfg.fact == FACT_BLOCKDIAG) {
      data(0, 0) = u[i];
      SSMBase iso;
      iso.fact = FACT_ISO;
      lp += iso.logpdf(obs.b[i], data);
    } else {
      for (size_t e = 0; e < data.a.size(); ++e) data.a[e] = u[e];
      lp += P.ops.logpdf(obs.b[i], data);
    }
  }
  return lp;
}

// MarkovSequence.sample (estimators_and_losses.py:233-273) of a smoother's
// posterior (the terminal marginal and K-1 backward conditionals, in solve()'s
// layouts) for one sample, with its standard normal draws given (the
// reference draws them from JAX keys): base (K x n), row K-1 for the
// terminal marginal, row k for the conditional k.  The reverse scan
// propagates the previous sample with cond.apply_flat and samples the result.
// Writes the K samples (K x mean_size).  Dense and isotropic; returns 2
// otherwise.
int pdqo_sample(const pdqo_cfg* cfg, const Real* params, int K, const Real* term_mean, const Real* term_chol,
                const Real* cA, const Real* cm, const Real* cL, const Real* ctl, const Real* cto, const Real* base,
                Real* out) {
  Problem P;
  P.setup(*cfg, params);
  if (P.cfg.fact == FACT_BLOCKDIAG) return 2;
  const int n = P.n, ms = mean_size(P), cs = chol_size(P);
  Normal x = sample_with(read_normal(P, term_mean, term_chol).b[0], base + (size_t)(K - 1) * n);
  std::memcpy(out + (size_t)(K - 1) * ms, x.m.a.data(), sizeof(Real) * ms);
  for (int k = K - 2; k >= 0; --k) {
    const BCond c = read_cond(P, cA + (size_t)k * n * n, cm + (size_t)k * ms, cL + (size_t)k * cs, ctl + (size_t)k * n,
                              cto + (size_t)k * n);
    const Normal pred = P.ops.apply_flat(c.b[0], x.m);
    x = sample_with(pred, base + (size_t)k * n);
    std::memcpy(out + (size_t)k * ms, x.m.a.data(), sizeof(Real) * ms);
  }
  return 0;
}

// One step attempt from the state (mean, chol) at t with step dt and its
// error power (test hook for the error estimators): writes the proposed
// mean (mean_size) and returns error_power.
Real pdqo_step_error_power(const pdqo_cfg* cfg, const Real* params, const Real* mean, const Real* chol, Real t, Real dt,
                           Real* prop_mean) {
  Problem P;
  P.setup(*cfg, params);
  Sol prev = solver_init(P, t, read_normal(P, mean, chol));
  const Sol prop = solver_step(P, prev, dt);
  write_normal(P, prop.u, prop_mean, nullptr);
  return estimate_error_power(P, prev, prop, dt);
}

// pdqo_lml_from_y0 with the calibration tape: mode 1 records the output
// scales of every solver_dynamic step into tape[0:tape_len], mode 2 replays
// them.  Returns NaN if the tape is too short.
Real pdqo_lml_from_y0_tape(const pdqo_cfg* cfg, const Real* params, const Real* y0, const Real* grid, int K, int every,
                           const Real* data, const Real* stds, int nstd, int average, Real* tape, int tape_len,
                           int mode) {
  g_tape_mode = mode == 1 ? TAPE_RECORD : TAPE_REPLAY;
  g_tape = tape;
  g_tape_len = tape_len < 0 ? 0 : (size_t)tape_len;
  g_tape_pos = 0;
  Real v;
  try {
    v = pdqo_lml_from_y0(cfg, params, y0, grid, K, every, data, stds, nstd, average);
  } catch (const std::runtime_error&) {
    v = std::numeric_limits<Real>::quiet_NaN();
  }
  g_tape_mode = TAPE_OFF;
  g_tape = nullptr;
  return v;
}

// Central finite-difference gradient of pdqo_lml_from_y0 with respect to
// (params[0:np], y0[0:d]); the plain-path gradient baseline.
Real pdqo_lml_grad_fd(const pdqo_cfg* cfg, const Real* params, int np, const Real* y0, const Real* grid, int K,
                        int every, const Real* data, const Real* stds, int nstd, int average, Real h,
                        Real* grad) {
  std::vector<Real> p(params, params + np), y(y0, y0 + cfg->d);
  const Real f0 = pdqo_lml_from_y0(cfg, p.data(), y.data(), grid, K, every, data, stds, nstd, average);
  for (int i = 0; i < np + cfg->d; ++i) {
    Real* x = i < np ? &p[i] : &y[i - np];
    const Real x0 = *x;
    *x = x0 + h;
    const Real fp = pdqo_lml_from_y0(cfg, p.data(), y.data(), grid, K, every, data, stds, nstd, average);
    *x = x0 - h;
    const Real fm = pdqo_lml_from_y0(cfg, p.data(), y.data(), grid, K, every, data, stds, nstd, average);
    *x = x0;
    grad[i] = (fp - fm) / (2.0 * h);
  }
  return f0;
}

// Save-every-step harness output; K_max bounds the number of rows written.
int pdqo_solve_save_every_step(const pdqo_cfg* cfg, const Real* params, const Real* tcoeffs, Real init_std,
                               Real t0, Real t1, int K_max, pdqo_out* out) {
  Problem P;
  P.setup(*cfg, params);
  BNormal init = prior_init(P, tcoeffs, init_std);
  Outputs o;
  solve_save_every_step(P, init, t0, t1, o);
  if ((int)o.sols.size() + 1 > K_max) return -(int)o.sols.size() - 1;
  Final F;
  finalize(P, o, F);
  write_outputs(P, o, F, out);
  return (int)F.t.size();
}

// The LML configuration's forward pass with its Markov-sequence outputs.
int pdqo_solve_fixed_grid_checkpoints(const pdqo_cfg* cfg, const Real* params, const Real* tcoeffs,
                                      Real init_std, const Real* grid, int K, int every, pdqo_out* out) {
  Problem P;
  P.setup(*cfg, params);
  BNormal init = prior_init(P, tcoeffs, init_std);
  Outputs o;
  solve_fixed_grid_checkpoints(P, init, grid, K, every, o);
  Final F;
  finalize(P, o, F);
  write_outputs(P, o, F, out);
  return (int)F.t.size();
}

// Batched adaptive solves for the CPU baseline: instance i has its own
// params, tcoeffs and save_at row; outputs only the terminal mean and the
// attempt/accept counts.  OpenMP over instances.
int pdqo_solve_save_at_batch(const pdqo_cfg* cfg, int n_inst, const Real* params, int n_params,
                             const Real* tcoeffs, const Real* save_at, int K, Real* term_mean, long* attempts,
                             long* accepted, int n_threads) {
  int err = 0;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
#endif
  for (int i = 0; i < n_inst; ++i) {
    Problem P;
    P.setup(*cfg, params + (size_t)i * n_params);
    BNormal init = prior_init(P, tcoeffs + (size_t)i * (cfg->nu + 1) * cfg->d, 0.0);
    Outputs o;
    solve_save_at(P, init, save_at + (size_t)i * K, K, o);
    if (o.status != ST_OK) {
      err |= 1;
      continue;
    }
    Final F;
    finalize(P, o, F);
    write_normal(P, F.u.back(), term_mean + (size_t)i * mean_size(P), nullptr);
    attempts[i] = o.attempts;
    accepted[i] = o.accepted;
  }
  (void)n_threads;
  return err;
}

}  // extern "C"
