// Cooperative persistent kernel for the dense and isotropic factorisations.
//
// Mapping: a workgroup is one wavefront (64 lanes) split into 64/G groups of
// G lanes; each group owns one IVP instance at a time and pulls the next one
// from a device work queue when it finishes (uneven adaptive step counts).
// Every matrix of an instance lives in LDS; the square-root algebra
// (Householder QR in LAPACK's sign convention, triangular solves, GEMMs) is
// executed cooperatively by the group's lanes in FP64.  A "covariance
// dimension" nc and "mean columns" mc describe both factorisations:
//   dense:      nc = (nu+1) d, mc = 1, ko = d observation rows
//   isotropic:  nc = nu+1,     mc = d, ko = 1
// (ssm_impl_dense.py:15-363, ssm_impl_isotropic.py:72-403).
#include "pdq_args.h"
#include "pdq_device.h"

namespace pdq {

// ---------------------------------------------------------------- context
#ifndef PDQ_HOST_EMU
// The dynamic LDS of every cooperative launch.  A pointer into it that reaches
// a function through memory (the context) or as an argument of a function
// that is not inlined is a generic address, and its loads and stores compile
// to flat instructions, each waited for on both the vector-memory and the LDS
// counters.  lds_ptr rebuilds such a pointer from this symbol (the same
// address), so that the compiler emits LDS instructions for it.  Only
// pointers into the launch's LDS may pass through it.
extern __shared__ double pdq_coop_lds[];
template <class T>
__device__ __forceinline__ T* lds_ptr(T* p) {
  const double* base = pdq_coop_lds;
  return reinterpret_cast<T*>(pdq_coop_lds + (reinterpret_cast<const double*>(p) - base));
}
#else
template <class T>
inline T* lds_ptr(T* p) {
  return p;
}
#endif


// Slot scalar indices.
enum { SC_T = 0, SC_OS = 1, SC_NSTEPS = 2, SC_MLE_RUN = 3, SC_MLE_N = 4, SC_IDENT = 5 };

template <int G>
struct Ctx {
  Grp<G> g;
  const KArgs* a;
  const Layout* L;
  int nc, mc, mcp, ko, N, d, dense;  // mcp: primal mean columns (mc includes tangents)
  double* sA_;   // shared A (nc x nc); the exponential prior's e^A of the current step
  double* sQ_;   // shared Q (nc x nc); the exponential prior's Gramian factor
  const double* sD_;  // exponential prior: the drift's last d rows (d x nc, shared)
  int expo;           // exponential prior (dense)
  double* ws;    // instance workspace base
  double* prm_;  // vector-field parameters (LDS copy)
  // probe set of the Monte Carlo Jacobian handler (accepted steps so far), set
  // by the solve loop before each step
  mutable uint32_t lin_set = 0;
  // every pointer the context holds is into LDS
  __device__ double* slot(int s) const { return lds_ptr(ws + s * L->s_size); }
  __device__ double* W(int off) const { return lds_ptr(ws + off); }
  __device__ double* sA() const { return lds_ptr(sA_); }
  __device__ double* sQ() const { return lds_ptr(sQ_); }
  __device__ const double* sD() const { return lds_ptr(sD_); }
  __device__ double* prm() const { return lds_ptr(prm_); }
};

// ------------------------------------------------------------- primitives
template <int G>
__device__ __forceinline__ void g_copy(const Grp<G>& g, double* dst, const double* src, int n) {
  for (int i = g.l; i < n; i += G) dst[i] = src[i];
}
template <int G>
__device__ __forceinline__ void g_fill(const Grp<G>& g, double* dst, double v, int n) {
  for (int i = g.l; i < n; i += G) dst[i] = v;
}

// In-place R-only Householder QR of the m x n matrix A (leading dim ld) in
// LAPACK dgeqr2 convention (backend/linalg.py:8-37 -> geqrf); on exit the
// upper triangle of rows 0..min(m,n)-1 holds R and the strict lower part is
// zeroed.
template <int G>
__device__ void g_qr_r(const Grp<G>& g, double* A, int m, int n, int ld) {
  const int k = m < n ? m : n;
  for (int i = 0; i < k; ++i) {
    if (m - i <= 1) break;
    double ss = 0.0;
    for (int r = i + 1 + g.l; r < m; r += G) {
      const double x = A[r * ld + i];
      ss += x * x;
    }
    ss = g.sum(ss);
    if (ss == 0.0) continue;
    const double alpha = A[i * ld + i];
    // beta != 0 and alpha - beta != 0 (ss > 0): the refined hardware
    // estimates stand in for hypot and the divisions, as in the lane kernel
    const double beta = -copysign(lane_sqrt_pos(fma(alpha, alpha, ss)), alpha);
    const double tau = lane_div_nz(beta - alpha, beta, lane_rcp_nz(beta));
    const double scal = lane_rcp_nz(alpha - beta);
    g.sync();
    for (int r = i + 1 + g.l; r < m; r += G) A[r * ld + i] *= scal;
    g.sync();
    for (int j = i + 1 + g.l; j < n; j += G) {
      double w = A[i * ld + j];
      for (int r = i + 1; r < m; ++r) w += A[r * ld + i] * A[r * ld + j];
      w *= tau;
      A[i * ld + j] -= w;
      for (int r = i + 1; r < m; ++r) A[r * ld + j] -= A[r * ld + i] * w;
    }
    if (g.l == 0) A[i * ld + i] = beta;
    g.sync();
  }
  // zero the Householder vectors (strict lower part) of the R rows
  for (int e = g.l; e < k * n; e += G) {
    const int i = e / n, j = e % n;
    if (j < i) A[i * ld + j] = 0.0;
  }
  g.sync();
}

// Singular value decomposition of the k x k upper-triangular R_Y (rows of R,
// leading dim ldr) by one-sided Jacobi rotations (Hestenes): on return U
// (k x k, row-major) holds R_Y V with mutually orthogonal columns, V (k x k)
// the accumulated rotations and s[j] the norm of column j of U.  Sweeps stop
// once every pair's cosine is below 1e-16 (at most 60).  One lane runs it.
__device__ inline void jacobi_svd(const double* R, int ldr, int k, double* U, double* V, double* s) {
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) {
      U[i * k + j] = R[i * ldr + j];
      V[i * k + j] = i == j ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < k; ++p)
      for (int q = p + 1; q < k; ++q) {
        double alpha = 0.0, beta = 0.0, gamma = 0.0;
        for (int i = 0; i < k; ++i) {
          alpha += U[i * k + p] * U[i * k + p];
          beta += U[i * k + q] * U[i * k + q];
          gamma += U[i * k + p] * U[i * k + q];
        }
        if (gamma == 0.0) continue;
        off = fmax(off, fabs(gamma) / sqrt(alpha * beta));
        // the rotation that makes columns p and q orthogonal (the smaller root of
        // t^2 + 2 zeta t - 1 = 0)
        const double zeta = (beta - alpha) / (2.0 * gamma);
        const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
        for (int i = 0; i < k; ++i) {
          const double up = U[i * k + p], uq = U[i * k + q];
          U[i * k + p] = cs * up - sn * uq;
          U[i * k + q] = sn * up + cs * uq;
          const double vp = V[i * k + p], vq = V[i * k + q];
          V[i * k + p] = cs * vp - sn * vq;
          V[i * k + q] = sn * vp + cs * vq;
        }
      }
    if (off < 1e-16) break;
  }
  for (int j = 0; j < k; ++j) {
    double ss = 0.0;
    for (int i = 0; i < k; ++i) ss += U[i * k + j] * U[i * k + j];
    s[j] = sqrt(ss);
  }
}

// G := solve(R_Y, R12)^T (cholesky_util.py:77-78): R_Y k x k upper (leading
// dim ldr), R12 k x n (same array, columns k..k+n-1), out n x k.
// mode 0: solve_triu (back substitution).
// mode 1: lstsq_svd (backend/linalg.py:60-61, jnp.linalg.lstsq): the
// minimum-norm least-squares solution, singular values at or below
// eps * k * s_max treated as zero; ws holds 2 k^2 + k doubles of the group's
// LDS.  Lane 0 decomposes R_Y, then each lane solves its columns.
template <int G>
__device__ void g_gain(const Grp<G>& g, const double* R, int ldr, int k, int n, double* out, int mode,
                       double* ws = nullptr) {
  if (mode == 1) {
    double* U = ws;
    double* V = ws + k * k;
    double* s = ws + 2 * k * k;
    if (g.l == 0) jacobi_svd(R, ldr, k, U, V, s);
    g.sync();
    double smax = 0.0;
    for (int j = 0; j < k; ++j) smax = fmax(smax, s[j]);
    const double cutoff = 2.220446049250313e-16 * (double)k * smax;
    for (int j = g.l; j < n; j += G) {
      for (int i = 0; i < k; ++i) out[j * k + i] = 0.0;
      for (int c = 0; c < k; ++c) {
        if (!(s[c] > cutoff)) continue;
        double ub = 0.0;
        for (int i = 0; i < k; ++i) ub += (U[i * k + c] / s[c]) * R[i * ldr + k + j];
        ub /= s[c];
        for (int i = 0; i < k; ++i) out[j * k + i] += V[i * k + c] * ub;
      }
    }
    return;
  }
  for (int j = g.l; j < n; j += G) {
    // back substitution for column j of R12, written straight into out
    for (int i = k - 1; i >= 0; --i) {
      double s = R[i * ldr + k + j];
      for (int c = i + 1; c < k; ++c) s -= R[i * ldr + c] * out[j * k + c];
      out[j * k + i] = s / R[i * ldr + i];
    }
  }
}

// ------------------------------------------------ exponential prior
// Z = X Y for row-major X (n x k) and Y (k x m), k ascending (jnp matmul);
// Z must not alias X or Y.
template <int G>
__device__ void g_gemm(const Grp<G>& g, const double* X, const double* Y, double* Z, int n, int k, int m) {
  for (int e = g.l; e < n * m; e += G) {
    const int i = e / m, j = e % m;
    double acc = 0.0;
    for (int q = 0; q < k; ++q) acc += X[i * k + q] * Y[q * m + j];
    Z[e] = acc;
  }
}

// jnp.linalg.solve (backend/linalg.py solve_lu): LU with partial pivoting of
// the n x n matrix M in place (the first largest |entry| of each column
// pivots, as LAPACK's idamax), the row swaps applied to the right-hand sides
// X1 (n x m1) and X2 (n x m2), then forward and back substitution.
template <int G>
__device__ void g_lu_solve(const Grp<G>& g, double* M, int n, double* X1, int m1, double* X2, int m2) {
  for (int k = 0; k < n; ++k) {
    double best = -1.0;
    int bi = k;
    for (int r = k + g.l; r < n; r += G) {
      const double v = fabs(M[r * n + k]);
      if (v > best) {
        best = v;
        bi = r;
      }
    }
    const double gbest = g.max(best);
    // the smallest row index among the lanes holding the maximum
    const double cand = g.max(best == gbest ? -(double)bi : -(double)n);
    int piv = (int)(-cand);
    piv = piv < k ? k : (piv >= n ? k : piv);
    if (piv != k) {
      for (int j = g.l; j < n; j += G) {
        const double t = M[k * n + j];
        M[k * n + j] = M[piv * n + j];
        M[piv * n + j] = t;
      }
      for (int j = g.l; j < m1; j += G) {
        const double t = X1[k * m1 + j];
        X1[k * m1 + j] = X1[piv * m1 + j];
        X1[piv * m1 + j] = t;
      }
      for (int j = g.l; j < m2; j += G) {
        const double t = X2[k * m2 + j];
        X2[k * m2 + j] = X2[piv * m2 + j];
        X2[piv * m2 + j] = t;
      }
    }
    g.sync();
    const double inv_p = 1.0 / M[k * n + k];
    for (int r = k + 1 + g.l; r < n; r += G) {
      const double l = M[r * n + k] * inv_p;
      M[r * n + k] = l;
      for (int j = k + 1; j < n; ++j) M[r * n + j] -= l * M[k * n + j];
    }
    g.sync();
  }
  for (int e = g.l; e < m1 + m2; e += G) {
    double* X = e < m1 ? X1 : X2;
    const int m = e < m1 ? m1 : m2, j = e < m1 ? e : e - m1;
    for (int i = 1; i < n; ++i) {
      double s = X[i * m + j];
      for (int q = 0; q < i; ++q) s -= M[i * n + q] * X[q * m + j];
      X[i * m + j] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = X[i * m + j];
      for (int q = i + 1; q < n; ++q) s -= M[i * n + q] * X[q * m + j];
      X[i * m + j] = s / M[i * n + i];
    }
  }
  g.sync();
}

// pade_and_legendre_9 (util/gram_util.py:261-331): Pade coefficients and the
// Legendre coefficient matrix.
__device__ __forceinline__ double pade9(int i) {
  constexpr double c[10] = {17643225600.0, 8821612800.0, 2075673600.0, 302702400.0, 30270240.0,
                            2162160.0,     110880.0,     3960.0,       90.0,        1.0};
  return c[i];
}
__device__ __forceinline__ double legendre9(int i, int j) {
  constexpr double c[10][10] = {{17643225600.0, 0, 605404800.0, 0, 4324320.0, 0, 7920.0, 0, 2.0, 0},
                                {0, 8821612800.0, 0, 155675520.0, 0, 617760.0, 0, 528.0, 0, 0},
                                {0, 0, 1470268800.0, 0, 15444000.0, 0, 34320.0, 0, 10.0, 0},
                                {0, 0, 0, 147026880.0, 0, 960960.0, 0, 1092.0, 0, 0},
                                {0, 0, 0, 0, 10501920.0, 0, 42120.0, 0, 18.0, 0},
                                {0, 0, 0, 0, 0, 583440.0, 0, 1320.0, 0, 0},
                                {0, 0, 0, 0, 0, 0, 26520.0, 0, 26.0, 0},
                                {0, 0, 0, 0, 0, 0, 0, 1020.0, 0, 0},
                                {0, 0, 0, 0, 0, 0, 0, 0, 34.0, 0},
                                {0, 0, 0, 0, 0, 0, 0, 0, 0, 1.0}};
  return c[i][j];
}

// DenseExponential.transition (ssm_impl_dense.py:405-423) for the step dt,
// given the preconditioner P, PI: e^{A_p} into sA and the lower-triangular
// Gramian factor U into sQ, with A_p = dt p^-1 A p and B_p = sqrt|dt| |p^-1| B
// (exp_gram_cholesky, util/gram_util.py:17-77: scaling by 2^-s, the
// degree-9 Pade / Legendre approximation, s doublings, non-negative
// diagonal).  The noise factor of the transition is then output_scale * U.
template <int G>
__device__ void exp_transition(const Ctx<G>& c, double dt, const double* P, const double* PI) {
  const Layout& L = *c.L;
  const int n = c.nc, d = c.d, N = c.N, m10 = 10 * d;
  double* eA = c.sA();
  double* U = c.sQ();
  double* Ap = c.W(L.w_xa);
  double* Bp = c.W(L.w_xB);
  double* A2 = c.W(L.w_xA2);
  double* A4 = c.W(L.w_xA4);
  double* A6 = c.W(L.w_xA6);
  double* A8 = c.W(L.w_xA8);
  double* T = c.W(L.w_xT);
  double* V = c.W(L.w_xV);
  double* VmU = c.W(L.w_xVmU);
  double* Pw = c.W(L.w_xP);
  double* PB = c.W(L.w_xPB);
  double* rhs = c.W(L.w_xrhs);
  double* tmp = c.W(L.w_xtmp);
  double* buf = c.W(L.w_xbuf);
  const double base = c.a->cfg.base_output_scale;
  for (int e = c.g.l; e < n * n; e += G) {
    const int i = e / n, j = e % n;
    const double a = i < n - d ? (j == i + d ? 1.0 : 0.0) : c.sD()[(i - (n - d)) * n + j];
    Ap[e] = ((dt * PI[i]) * a) * P[j];
  }
  for (int e = c.g.l; e < n * d; e += G) {
    const int i = e / d, k = e % d;
    Bp[e] = (sqrt(fabs(dt)) * fabs(PI[i])) * (i == (N - 1) * d + k ? base : 0.0);
  }
  c.g.sync();
  // _exp_gram_cholesky_init: the number of doublings from the 1-norm
  double colmax = 0.0;
  for (int j = c.g.l; j < n; j += G) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += fabs(Ap[i * n + j]);
    colmax = fmax(colmax, s);
  }
  const double a_l1 = c.g.max(colmax);
  const double s1 = log2(a_l1 / 0.41098379928173995), s2 = log2((double)(n - 1) / 9.0);
  // at most 64 doublings (a 1-norm up to 0.41 2^64): a larger or non-finite
  // norm gives a non-finite transition either way, and the loop stays bounded
  double num = fmax(0.0, ceil(fmax(s1, s2)));
  if (!(num <= 64.0)) num = 64.0;
  const int nd = (int)num;
  const double two = ldexp(1.0, nd), sq = sqrt(two);
  for (int e = c.g.l; e < n * n; e += G) Ap[e] = (1.0 / two) * Ap[e];
  for (int e = c.g.l; e < n * d; e += G) Bp[e] = Bp[e] / sq;
  c.g.sync();
  g_gemm(c.g, Ap, Ap, A2, n, n, n);
  c.g.sync();
  g_gemm(c.g, A2, A2, A4, n, n, n);
  c.g.sync();
  g_gemm(c.g, A4, A2, A6, n, n, n);
  c.g.sync();
  g_gemm(c.g, A6, A2, A8, n, n, n);
  c.g.sync();
  for (int e = c.g.l; e < n * n; e += G) {
    const double id = (e / n == e % n) ? 1.0 : 0.0;
    T[e] = (((pade9(9) * A8[e] + pade9(7) * A6[e]) + pade9(5) * A4[e]) + pade9(3) * A2[e]) + pade9(1) * id;
    V[e] = (((pade9(8) * A8[e] + pade9(6) * A6[e]) + pade9(4) * A4[e]) + pade9(2) * A2[e]) + pade9(0) * id;
  }
  c.g.sync();
  g_gemm(c.g, Ap, T, Pw, n, n, n);  // U of the Pade approximant
  c.g.sync();
  for (int e = c.g.l; e < n * n; e += G) {
    VmU[e] = V[e] - Pw[e];
    eA[e] = V[e] + Pw[e];
  }
  // Legendre terms: blocks 2i (even, L_even[i]) and 2i+1 (odd, L_odd[i]) of rhs
  g_gemm(c.g, A2, Bp, PB, n, n, d);
  c.g.sync();
  for (int e = c.g.l; e < n * m10; e += G) {
    const int r = e / m10, col = e % m10, blk = col / d, k = col % d;
    const double pb = PB[r * d + k], b = Bp[r * d + k];
    double v = 0.0;
    if (blk == 0) v = legendre9(0, 2) * pb + legendre9(0, 0) * b;
    else if (blk == 1) v = legendre9(1, 3) * pb + legendre9(1, 1) * b;
    else if (blk == 2) v = legendre9(2, 2) * pb;
    else if (blk == 3) v = legendre9(3, 3) * pb;
    rhs[e] = v;
  }
  for (int e = c.g.l; e < n * n; e += G) Pw[e] = A2[e];
  c.g.sync();
  for (int kk = 2; kk <= 4; ++kk) {
    g_gemm(c.g, A2, Pw, T, n, n, n);
    c.g.sync();
    for (int e = c.g.l; e < n * n; e += G) Pw[e] = T[e];
    c.g.sync();
    g_gemm(c.g, Pw, Bp, PB, n, n, d);
    c.g.sync();
    for (int e = c.g.l; e < n * m10; e += G) {
      const int r = e / m10, col = e % m10, blk = col / d, k = col % d;
      const int i = blk / 2;
      const double coef = (blk % 2 == 0) ? legendre9(2 * i, 2 * kk) : legendre9(2 * i + 1, 2 * kk + 1);
      rhs[e] = rhs[e] + coef * PB[r * d + k];
    }
    c.g.sync();
  }
  // L_odd = A L_odd, then every block / sqrt(2 b + 1)
  for (int e = c.g.l; e < n * 5 * d; e += G) {
    const int r = e / (5 * d), col = e % (5 * d), i = col / d, k = col % d;
    double acc = 0.0;
    for (int q = 0; q < n; ++q) acc += Ap[r * n + q] * rhs[q * m10 + (2 * i + 1) * d + k];
    tmp[e] = acc;
  }
  c.g.sync();
  for (int e = c.g.l; e < n * m10; e += G) {
    const int r = e / m10, col = e % m10, blk = col / d, k = col % d;
    const double v = (blk % 2 == 1) ? tmp[r * 5 * d + (blk / 2) * d + k] : rhs[e];
    rhs[e] = v / sqrt((double)(2 * blk + 1));
  }
  c.g.sync();
  g_lu_solve(c.g, VmU, n, eA, n, rhs, m10);
  // qr_r(L^T)^T: the Gramian factor of the scaled problem
  for (int e = c.g.l; e < m10 * n; e += G) {
    const int i = e / n, j = e % n;
    buf[e] = rhs[j * m10 + i];
  }
  c.g.sync();
  g_qr_r(c.g, buf, m10, n, n);
  const int kr = m10 < n ? m10 : n;
  for (int e = c.g.l; e < n * n; e += G) {
    const int j = e / n, i = e % n;  // U[j][i] = R[i][j]
    U[e] = i < kr ? buf[i * n + j] : 0.0;
  }
  c.g.sync();
  // _exp_gram_cholesky_double
  for (int it = 0; it < nd; ++it) {
    g_gemm(c.g, eA, U, T, n, n, n);
    c.g.sync();
    for (int e = c.g.l; e < n * n; e += G) {
      const int r = e / n, col = e % n;
      buf[r * n + col] = U[col * n + r];
      buf[(n + r) * n + col] = T[col * n + r];
    }
    c.g.sync();
    g_qr_r(c.g, buf, 2 * n, n, n);
    for (int e = c.g.l; e < n * n; e += G) {
      const int j = e / n, i = e % n;
      U[e] = buf[i * n + j];
    }
    g_gemm(c.g, eA, eA, T, n, n, n);
    c.g.sync();
    for (int e = c.g.l; e < n * n; e += G) eA[e] = T[e];
    c.g.sync();
  }
  // non-negative diagonal (column signs)
  for (int j = c.g.l; j < n; j += G) {
    const double djj = U[j * n + j];
    const double sgn = djj < 0.0 ? -1.0 : 1.0;
    for (int i = 0; i < n; ++i) U[i * n + j] = U[i * n + j] * sgn;
  }
  c.g.sync();
}

// ------------------------------------------------ IWP prior and transitions
// preconditioner_taylor (utilities.py:74-84) expanded to nc entries; with the
// exponential prior also that step's transition (exp_transition).
template <int G>
__device__ void precon(const Ctx<G>& c, double dt, double* P_, double* PI_) {
  double* const P = lds_ptr(P_);
  double* const PI = lds_ptr(PI_);
  for (int i = c.g.l; i < c.nc; i += G) {
    const int deriv = c.dense ? i / c.d : i;
    const int pw = c.N - 1 - deriv;
    const double s = factorial_d(pw);
    P[i] = pow(dt, (double)pw) / s;
    PI[i] = pow(dt, -(double)pw) * s;
  }
  if (c.expo) {
    c.g.sync();
    exp_transition(c, dt, P, PI);
  }
}

// Scalar factor of the transition noise: dense sqrt|dt| * s (Q holds the base
// scale), isotropic (sqrt|dt| * base) * s (ssm_impl_dense.py:360-362,
// ssm_impl_isotropic.py:375-376); the exponential prior's factor already
// holds sqrt|dt| and the base scale (ssm_impl_dense.py:421-423).
template <int G>
__device__ __forceinline__ double noise_scale(const Ctx<G>& c, double dt, double s) {
  if (c.expo) return s;
  if (c.dense) return sqrt(fabs(dt)) * s;
  return sqrt(fabs(dt)) * c.a->cfg.base_output_scale * s;
}

// apply_flat of the transition to a mean x (ssm_impl_dense.py:18-22,
// ssm_impl_isotropic.py:75-79): out = p * (A (p^-1 * x)) (dense) or
// (p * A) (p^-1 * x) (isotropic); the noise mean is zero.
template <int G>
__device__ void apply_flat_mean(const Ctx<G>& c, const double* P, const double* PI, const double* x, double* out) {
  const int nc = c.nc, mc = c.mc;
  for (int e = c.g.l; e < nc * mc; e += G) {
    const int i = e / mc, col = e % mc;
    double acc = 0.0;
    if (c.dense) {
      for (int k = 0; k < nc; ++k) acc += c.sA()[i * nc + k] * (PI[k] * x[k * mc + col]);
      out[e] = P[i] * (acc + 0.0);
    } else {
      for (int k = 0; k < nc; ++k) acc += (P[i] * c.sA()[i * nc + k]) * (PI[k] * x[k * mc + col]);
      out[e] = acc + 0.0;
    }
  }
}

// chol of apply_flat: |p| * (s Q)
template <int G>
__device__ void apply_flat_chol(const Ctx<G>& c, const double* P, double s, double* out) {
  const int nc = c.nc;
  for (int e = c.g.l; e < nc * nc; e += G) {
    const int i = e / nc;
    out[e] = fabs(P[i]) * (s * c.sQ()[e]);
  }
}

// marginalise of the transition (ssm_impl_dense.py:24-33,
// ssm_impl_isotropic.py:81-89): (Xm, XL) -> (Ym, YL).  big: 2nc x nc.
template <int G>
__device__ void marginalise_prior(const Ctx<G>& c, const double* P, const double* PI, double s, const double* Xm,
                                  const double* XL, double* Ym, double* YL, double* big) {
  const int nc = c.nc, mc = c.mc;
  // rows 0..nc-1: (A (p^-1 * L))^T ; rows nc..2nc-1: (s Q)^T
  for (int e = c.g.l; e < nc * nc; e += G) {
    const int r = e / nc, col = e % nc;  // big[r][col] = (A chol_l)[col][r]
    double acc = 0.0;
    for (int k = 0; k < nc; ++k) acc += c.sA()[col * nc + k] * (PI[k] * XL[k * nc + r]);
    big[r * nc + col] = acc;
    big[(nc + r) * nc + col] = s * c.sQ()[col * nc + r];
  }
  for (int e = c.g.l; e < nc * mc; e += G) {
    const int i = e / mc, col = e % mc;
    double acc = 0.0;
    for (int k = 0; k < nc; ++k) acc += c.sA()[i * nc + k] * (PI[k] * Xm[k * mc + col]);
    Ym[e] = P[i] * (acc + 0.0);
  }
  c.g.sync();
  g_qr_r(c.g, big, 2 * nc, nc, nc);
  for (int e = c.g.l; e < nc * nc; e += G) {
    const int i = e / nc, j = e % nc;
    YL[e] = fabs(P[i]) * big[j * nc + i];
  }
}

// revert of the transition (ssm_impl_dense.py:51-77, isotropic :107-133,
// cholesky_util.py:27-82): X -> observed (Ym, YL) and backward conditional.
template <int G>
__device__ void revert_prior(const Ctx<G>& c, const double* P, const double* PI, double s, const double* Xm,
                             const double* XL, double* Ym, double* YL, double* bA, double* bm, double* bL, double* btl,
                             double* bto, double* big, double* tmp) {
  const int nc = c.nc, mc = c.mc, n2 = 2 * nc;
  // [[R_YX, 0], [R_X_F, R_X]]; R_YX = (sQ)^T, R_X_F = (A chol_l)^T, R_X = chol_l^T
  for (int e = c.g.l; e < nc * nc; e += G) {
    const int r = e / nc, col = e % nc;
    big[r * n2 + col] = s * c.sQ()[col * nc + r];
    big[r * n2 + nc + col] = 0.0;
    double acc = 0.0;
    for (int k = 0; k < nc; ++k) acc += c.sA()[col * nc + k] * (fabs(PI[k]) * XL[k * nc + r]);
    big[(nc + r) * n2 + col] = acc;
    big[(nc + r) * n2 + nc + col] = fabs(PI[col]) * XL[col * nc + r];
  }
  // mean_observed = A (p^-1 * m)
  for (int e = c.g.l; e < nc * mc; e += G) {
    const int i = e / mc, col = e % mc;
    double acc = 0.0;
    for (int k = 0; k < nc; ++k) acc += c.sA()[i * nc + k] * (PI[k] * Xm[k * mc + col]);
    tmp[e] = acc + 0.0;
  }
  c.g.sync();
  g_qr_r(c.g, big, n2, n2, n2);
  g_gain(c.g, big, n2, nc, nc, bA, 0);
  c.g.sync();
  for (int e = c.g.l; e < nc * mc; e += G) {
    const int i = e / mc, col = e % mc;
    double acc = 0.0;
    for (int k = 0; k < nc; ++k) acc += bA[i * nc + k] * tmp[k * mc + col];
    bm[e] = PI[i] * Xm[e] - acc;
    Ym[e] = P[i] * tmp[e];
  }
  for (int e = c.g.l; e < nc * nc; e += G) {
    const int i = e / nc, j = e % nc;
    bL[e] = big[(nc + j) * n2 + nc + i];
    YL[e] = fabs(P[i]) * big[j * n2 + i];
  }
  for (int i = c.g.l; i < nc; i += G) {
    btl[i] = 1.0 / P[i];
    bto[i] = 1.0 / PI[i];
  }
}

// merge (ssm_impl_dense.py:35-49, isotropic :91-105): out = self o other.
// self_ident: self is the identity conditional (exact shortcut, see below).
template <int G>
__device__ void merge(const Ctx<G>& c, bool self_ident, const double* sA, const double* sm, const double* sL,
                      const double* stl, const double* sto, const double* oA, const double* om, const double* oL,
                      const double* otl, const double* oto, double* rA, double* rm, double* rL, double* rtl,
                      double* rto, double* big) {
  const int nc = c.nc, mc = c.mc;
  if (self_ident) {
    // I (T o A2) == T o A2 and qr_r of the upper-triangular (|T| o L2)^T
    // stacked on zeros is that matrix itself (dlarfg skips zero
    // sub-columns), so the general formula reduces exactly to this.
    for (int e = c.g.l; e < nc * nc; e += G) {
      const int i = e / nc;
      const double T = 1.0 * oto[i];
      rA[e] = T * oA[e];
      rL[e] = fabs(T) * oL[e];
    }
    for (int e = c.g.l; e < nc * mc; e += G) {
      const int i = e / mc;
      rm[e] = (1.0 * oto[i]) * om[e] + 0.0;
    }
    for (int i = c.g.l; i < nc; i += G) {
      rtl[i] = otl[i];
      rto[i] = 1.0;
    }
    c.g.sync();
    return;
  }
  for (int e = c.g.l; e < nc * nc; e += G) {
    const int i = e / nc, j = e % nc;
    double acc = 0.0, acc2 = 0.0;
    for (int k = 0; k < nc; ++k) {
      const double T = stl[k] * oto[k];
      acc += sA[i * nc + k] * (T * oA[k * nc + j]);
      acc2 += sA[i * nc + k] * (fabs(T) * oL[k * nc + j]);
    }
    rA[e] = acc;
    big[j * nc + i] = acc2;             // R1 = (A1 (|T| o C2))^T
    big[(nc + j) * nc + i] = sL[e];     // R2 = C1^T
  }
  for (int e = c.g.l; e < nc * mc; e += G) {
    const int i = e / mc, col = e % mc;
    double acc = 0.0;
    for (int k = 0; k < nc; ++k) acc += sA[i * nc + k] * ((stl[k] * oto[k]) * om[k * mc + col]);
    rm[e] = acc + sm[e];
  }
  c.g.sync();
  g_qr_r(c.g, big, 2 * nc, nc, nc);
  for (int e = c.g.l; e < nc * nc; e += G) {
    const int i = e / nc, j = e % nc;
    rL[e] = big[j * nc + i];
  }
  for (int i = c.g.l; i < nc; i += G) {
    rtl[i] = otl[i];
    rto[i] = sto[i];
  }
  c.g.sync();
}

// marginalise of a general (backward) conditional; used by the smoother's
// backward pass (estimators_and_losses.py:156-178).
template <int G>
__device__ void marginalise_cond(const Ctx<G>& c, const double* cA, const double* cm, const double* cL,
                                 const double* ctl, const double* cto, const double* Xm, const double* XL, double* Ym,
                                 double* YL, double* big) {
  const int nc = c.nc, mc = c.mc;
  for (int e = c.g.l; e < nc * nc; e += G) {
    const int r = e / nc, col = e % nc;
    double acc = 0.0;
    for (int k = 0; k < nc; ++k) acc += cA[col * nc + k] * (ctl[k] * XL[k * nc + r]);
    big[r * nc + col] = acc;
    big[(nc + r) * nc + col] = cL[col * nc + r];
  }
  for (int e = c.g.l; e < nc * mc; e += G) {
    const int i = e / mc, col = e % mc;
    double acc = 0.0;
    for (int k = 0; k < nc; ++k) acc += cA[i * nc + k] * (ctl[k] * Xm[k * mc + col]);
    Ym[e] = cto[i] * (acc + cm[e]);
  }
  c.g.sync();
  g_qr_r(c.g, big, 2 * nc, nc, nc);
  for (int e = c.g.l; e < nc * nc; e += G) {
    const int i = e / nc, j = e % nc;
    YL[e] = fabs(cto[i]) * big[j * nc + i];
  }
}

// ------------------------------------------------------- linearisation
// Row selector of ts0 (derivative 1) as an index into the state.
template <int G>
__device__ __forceinline__ int sel_row(const Ctx<G>& c, int obs) {
  return c.dense ? c.d + obs : 1;
}

// ts0 / ts1 linearisation at the mean Xm (ssm_impl_dense.py:237-334,
// ssm_impl_isotropic.py:298-317): writes b (ko x mc) and, for ts1, J (ko x nc).
// With tangent columns (MODE_LML) b also carries d b / d theta.
template <int G>
__device__ void linearize(const Ctx<G>& c, const double* Xm_, double t, double* b_, double* J_, double* y0buf_,
                          uint32_t set_offset = 0) {
  const double* const Xm = lds_ptr(Xm_);
  double* const b = lds_ptr(b_);
  double* const J = lds_ptr(J_);
  double* const y0buf = lds_ptr(y0buf_);
  const int d = c.d, mc = c.mc;
  const int id = c.a->cfg.vector_field;
  // gather u(t) (zeroth Taylor coefficient) and, with tangents, its tangents
  const int ntan = c.mcp == c.mc ? 0 : c.a->n_tan;
  for (int k = c.g.l; k < d * (1 + ntan); k += G) {
    const int q = k / d, j = k % d;
    y0buf[k] = c.dense ? Xm[j * mc + q] : Xm[q * d + j];
  }
  c.g.sync();
  const double* y = y0buf;
  if (c.a->cfg.constraint == PDQ_TS0) {
    for (int k = c.g.l; k < d * (1 + ntan); k += G) {
      const int q = k / d, j = k % d;
      double v;
      if (q == 0) {
        v = -vf_component(id, d, y, c.prm(), t, j);
      } else {
        double acc = 0.0;
        for (int jj = 0; jj < d; ++jj) acc += vf_jac(id, d, y, c.prm(), j, jj) * y[q * d + jj];
        const int qq = q - 1;
        if (qq < c.a->cfg.n_params) acc += vf_jac_param(id, d, y, c.prm(), j, qq);
        v = -acc;
      }
      if (c.dense)
        b[j * mc + q] = v;
      else
        b[q * d + j] = v;
    }
  } else if (!c.dense) {
    // ts1 (isotropic, IsotropicResidual.linearize, ssm_impl_isotropic.py:321-355,
    // with jacobian_materialize's exact trace, jacobians.py:100-106): the
    // 1 x nc operator J = [h0, 1, 0...] with h0 = -tr(J_f) / d, and
    // b = r(m) - J m for the residual r = m_1 - f(m_0).  With
    // jacobian_monte_carlo_rev / _fwd (:230-248, :151-169) tr J_f is the probe
    // mean of v^T J_f v.
    double tr = 0.0;
    if (c.a->cfg.jacobian != PDQ_JAC_EXACT) {
      const bool fwd = c.a->cfg.jacobian == PDQ_JAC_MONTE_CARLO_FWD;
      for (int j = 0; j < d; ++j)
        tr += -vf_mc_diag(id, d, y, c.prm(), j, c.a->cfg.jac_seed, c.lin_set + set_offset, c.a->cfg.jac_probes, fwd);
    } else {
      for (int j = 0; j < d; ++j) tr += -vf_jac(id, d, y, c.prm(), j, j);
    }
    const double h0 = tr / (double)d;
    for (int k = c.g.l; k < c.nc; k += G) J[k] = k == 0 ? h0 : (k == 1 ? 1.0 : 0.0);
    for (int j = c.g.l; j < d; j += G) {
      const double r = Xm[d + j] - vf_component(id, d, y, c.prm(), t, j);
      b[j] = r - (h0 * Xm[j] + Xm[d + j]);
    }
  } else {
    // ts1 (dense): J = [-J_f, I, 0...], b = (m_1 - f(m_0)) - J m
    const int nc = c.nc;
    for (int e = c.g.l; e < d * nc; e += G) {
      const int i = e / nc, j = e % nc;
      double v = 0.0;
      if (j < d) v = -vf_jac(id, d, y, c.prm(), i, j);
      else if (j == d + i) v = 1.0;
      J[e] = v;
    }
    c.g.sync();
    for (int i = c.g.l; i < d; i += G) {
      const double f = vf_component(id, d, y, c.prm(), t, i);
      double jm = 0.0;
      for (int k = 0; k < nc; ++k) jm += J[i * nc + k] * Xm[k * mc];
      b[i * mc] = (Xm[(d + i) * mc] - f) - jm;
    }
  }
  c.g.sync();
}

// (H L) entry (obs row o, column r) for ts0 / ts1.
template <int G>
__device__ __forceinline__ double HL(const Ctx<G>& c, const double* J, const double* XL, int o, int r) {
  if (c.a->cfg.constraint == PDQ_TS0) return XL[sel_row(c, o) * c.nc + r];
  double acc = 0.0;
  for (int k = 0; k < c.nc; ++k) acc += J[o * c.nc + k] * XL[k * c.nc + r];
  return acc;
}
template <int G>
__device__ __forceinline__ double Hm(const Ctx<G>& c, const double* J, const double* Xm, int o, int col) {
  if (c.a->cfg.constraint == PDQ_TS0) return Xm[sel_row(c, o) * c.mc + col];
  double acc = 0.0;
  for (int k = 0; k < c.nc; ++k) acc += J[o * c.nc + k] * Xm[k * c.mc + col];
  return acc;
}

// fx.marginalise(u) -> observed (om ko x mc, oL ko x ko).  big: (nc+ko) x ko.
template <int G>
__device__ void marginalise_fx(const Ctx<G>& c, const double* b, const double* J, const double* Xm, const double* XL,
                               double* om, double* oL, double* big) {
  const int nc = c.nc, ko = c.ko, mc = c.mc;
  const double damp = c.a->cfg.damp;
  for (int e = c.g.l; e < (nc + ko) * ko; e += G) {
    const int r = e / ko, o = e % ko;
    big[e] = r < nc ? HL(c, J, XL, o, r) : (r - nc == o ? damp : 0.0);
  }
  for (int e = c.g.l; e < ko * mc; e += G) {
    const int o = e / mc, col = e % mc;
    om[e] = Hm(c, J, Xm, o, col) + b[e];
  }
  c.g.sync();
  g_qr_r(c.g, big, nc + ko, ko, ko);
  for (int e = c.g.l; e < ko * ko; e += G) {
    const int i = e / ko, j = e % ko;
    oL[e] = big[j * ko + i];
  }
  c.g.sync();
}

// residual_whitened_rms against zero data (ssm_impl_dense.py:156-160,
// isotropic :203-207), primal columns only.  Every lane returns the value.
template <int G>
__device__ double rms_zero(const Ctx<G>& c, const double* om, const double* oL) {
  const int ko = c.ko, mc = c.mc, mcp = c.mcp;
  double* w = c.W(c.L->w_tri);
  double ss = 0.0;
  for (int col = c.g.l; col < mcp; col += G) {
    for (int i = 0; i < ko; ++i) {
      double s = c.dense ? (0.0 - om[i * mc + col]) : om[i * mc + col];
      for (int j = 0; j < i; ++j) s -= oL[i * ko + j] * w[col * ko + j];
      const double wi = s / oL[i * ko + i];
      w[col * ko + i] = wi;
      ss += wi * wi;
    }
  }
  ss = c.g.sum(ss);
  c.g.sync();
  return sqrt(ss) / sqrt((double)(ko * mcp));
}

// fx.revert(pred) and apply_flat(data) (ssm_impl_api.py:112-132):
// (Pm, PL) -> (Xm, XL); optionally returns the observed normal.  big:
// (ko+nc)^2, gain: nc x ko, mo: ko x mc.
template <int G>
__device__ void update(const Ctx<G>& c, const double* b, const double* J, const double* Pm, const double* PL,
                       double* Xm, double* XL, double* big, double* gain, double* mo, double* oL_out) {
  const int nc = c.nc, ko = c.ko, mc = c.mc, n2 = ko + nc;
  const double damp = c.a->cfg.damp;
  for (int e = c.g.l; e < n2 * n2; e += G) {
    const int r = e / n2, col = e % n2;
    double v;
    if (r < ko) v = col < ko ? (r == col ? damp : 0.0) : 0.0;
    else v = col < ko ? HL(c, J, PL, col, r - ko) : PL[(col - ko) * nc + (r - ko)];
    big[e] = v;
  }
  for (int e = c.g.l; e < ko * mc; e += G) {
    const int o = e / mc, col = e % mc;
    mo[e] = Hm(c, J, Pm, o, col) + b[e];
  }
  c.g.sync();
  g_qr_r(c.g, big, n2, n2, n2);
  g_gain(c.g, big, n2, ko, nc, gain, 0);
  c.g.sync();
  for (int e = c.g.l; e < nc * mc; e += G) {
    const int i = e / mc, col = e % mc;
    double acc = 0.0;
    for (int o = 0; o < ko; ++o) acc += gain[i * ko + o] * mo[o * mc + col];
    Xm[e] = Pm[e] - acc;
  }
  for (int e = c.g.l; e < nc * nc; e += G) {
    const int i = e / nc, j = e % nc;
    XL[e] = big[(ko + j) * n2 + ko + i];
  }
  if (oL_out)
    for (int e = c.g.l; e < ko * ko; e += G) {
      const int i = e / ko, j = e % ko;
      oL_out[e] = big[j * n2 + i];
    }
  c.g.sync();
}

// ------------------------------------------------------------- error norm
// error_residual_std.estimate_error_norm (solvers.py:925-996) given the
// observed marginal of the error model and its local scale sigma.
template <int G>
__device__ double error_power(const Ctx<G>& c, const double* oL_, double sigma, const double* prev_m_,
                              const double* prop_m_, double dt) {
  const double* const oL = lds_ptr(oL_);
  const double* const prev_m = lds_ptr(prev_m_);
  const double* const prop_m = lds_ptr(prop_m_);
  const int ko = c.ko, d = c.d;
  const pdq_config& cfg = c.a->cfg;
  // std of rescaled observed rows: dense per-row qr_r (|beta| of dlarfg),
  // isotropic vector_norm
  auto err = [&](int r) {
    double ss = 0.0;
    for (int j = 1; j < ko; ++j) {
      const double x = sigma * oL[r * ko + j];
      ss += x * x;
    }
    const double alpha = sigma * oL[r * ko];
    if (c.dense) return ss == 0.0 ? fabs(alpha) : fabs(hypot(alpha, sqrt(ss)));
    return sqrt(alpha * alpha + ss);
  };
  // n = residual_order - 1 = 1, one more for error_per_unit_step
  const int nord = cfg.error_per_unit_step ? 2 : 1;
  const double fact = nord == 1 ? 1.0 : 2.0;
  double norm;
  if (cfg.error_norm == PDQ_NORM_SCALE_THEN_RMS) {
    double ss = 0.0;
    for (int k = 0; k < d; ++k) {
      const double u0 = c.dense ? prev_m[k * c.mc] : prev_m[k];
      const double u1 = c.dense ? prop_m[k * c.mc] : prop_m[k];
      const double ref = fmax(fabs(u0), fabs(u1));
      const double e = err(ko == 1 ? 0 : k) * pow(dt, (double)nord) / fact;
      const double rr = e / (cfg.atol + cfg.rtol * fabs(ref));
      ss += rr * rr;
    }
    norm = sqrt(ss) / sqrt((double)d);
  } else {
    double sa = 0.0, sr = 0.0;
    for (int r = 0; r < ko; ++r) {
      const double e = err(r) * pow(dt, (double)nord) / fact;
      sa += e * e;
    }
    for (int k = 0; k < d; ++k) {
      const double u0 = c.dense ? prev_m[k * c.mc] : prev_m[k];
      const double u1 = c.dense ? prop_m[k * c.mc] : prop_m[k];
      const double ref = fmax(fabs(u0), fabs(u1));
      sr += ref * ref;
    }
    norm = (sqrt(sa) / sqrt((double)ko)) / (cfg.atol + cfg.rtol * (sqrt(sr) / sqrt((double)d)));
  }
  return pow(norm, -1.0 / (double)c.N);
}

// The std of row r of a factor (n columns): dense per-row qr_r (|beta| of
// dlarfg, ssm_impl_dense.py:141-150), isotropic vector_norm (:189-197).
template <int G>
__device__ __forceinline__ double row_std(const Ctx<G>& c, const double* L, int r, int n) {
  double ss = 0.0;
  for (int j = 1; j < n; ++j) ss += L[r * n + j] * L[r * n + j];
  const double alpha = L[r * n];
  if (c.dense) return ss == 0.0 ? fabs(alpha) : fabs(hypot(alpha, sqrt(ss)));
  return sqrt(alpha * alpha + ss);
}

// error_state_std (solvers.py:1041-1098): condition the output-scale-1
// extrapolation (Um, UL) on the linearisation (b, J); the errors are the
// calibrated std of derivative idx of the result (d values dense, one
// isotropic), written to errs.  Xm, XL, mo, oL are scratch.
template <int G>
__device__ void state_std_errors(const Ctx<G>& c, const double* b_, const double* J_, const double* Um_,
                                 const double* UL_, double* Xm_, double* XL_, double* big_, double* gain_, double* mo_,
                                 double* oL_, double* errs_) {
  const double *const b = lds_ptr(b_), *const J = lds_ptr(J_), *const Um = lds_ptr(Um_), *const UL = lds_ptr(UL_);
  double *const Xm = lds_ptr(Xm_), *const XL = lds_ptr(XL_), *const big = lds_ptr(big_), *const gain = lds_ptr(gain_);
  double *const mo = lds_ptr(mo_), *const oL = lds_ptr(oL_), *const errs = lds_ptr(errs_);
  update(c, b, J, Um, UL, Xm, XL, big, gain, mo, oL);
  const double os = rms_zero(c, mo, oL);
  const int idx = c.a->cfg.error_derivative_idx;
  const int ne = c.dense ? c.d : 1;
  for (int k = c.g.l; k < ne; k += G) errs[k] = os * row_std(c, XL, c.dense ? idx * c.d + k : idx, c.nc);
  c.g.sync();
}

// The error power of error_state_std: errs (d values dense, one isotropic) at
// order n = idx (+1 per unit step), referenced to derivative idx.
template <int G>
__device__ double error_power_state(const Ctx<G>& c, const double* errs_, const double* prev_m_,
                                    const double* prop_m_, double dt) {
  const double* const errs = lds_ptr(errs_);
  const double* const prev_m = lds_ptr(prev_m_);
  const double* const prop_m = lds_ptr(prop_m_);
  const pdq_config& cfg = c.a->cfg;
  const int d = c.d, idx = cfg.error_derivative_idx, ne = c.dense ? d : 1;
  const int nord = idx + (cfg.error_per_unit_step ? 1 : 0);
  const double scale = pow(dt, (double)nord) / factorial_d(nord);
  auto ref = [&](int k) {
    const double u0 = c.dense ? prev_m[(idx * d + k) * c.mc] : prev_m[idx * c.mc + k];
    const double u1 = c.dense ? prop_m[(idx * d + k) * c.mc] : prop_m[idx * c.mc + k];
    return fmax(fabs(u0), fabs(u1));
  };
  double norm;
  if (cfg.error_norm == PDQ_NORM_SCALE_THEN_RMS) {
    double ss = 0.0;
    for (int k = 0; k < d; ++k) {
      const double e = errs[ne == 1 ? 0 : k] * scale;
      const double rr = e / (cfg.atol + cfg.rtol * fabs(ref(k)));
      ss += rr * rr;
    }
    norm = sqrt(ss) / sqrt((double)d);
  } else {
    double sa = 0.0, sr = 0.0;
    for (int k = 0; k < ne; ++k) {
      const double e = errs[k] * scale;
      sa += e * e;
    }
    for (int k = 0; k < d; ++k) sr += ref(k) * ref(k);
    norm = (sqrt(sa) / sqrt((double)ne)) / (cfg.atol + cfg.rtol * (sqrt(sr) / sqrt((double)d)));
  }
  return pow(norm, -1.0 / (double)c.N);
}

// controllers.py:66-84 / :24-63
__device__ __forceinline__ double control_apply(const pdq_config& cfg, double dt, double& cstate, double ep) {
  if (cfg.control == PDQ_CONTROL_INTEGRAL) {
    const double ratio = cfg.safety * ep;
    return fmax(cfg.factor_min, fmin(ratio, cfg.factor_max)) * dt;
  }
  const double gi = pow(ep, cfg.exponent_integral);
  const double gp = pow(ep / cstate, cfg.exponent_proportional);
  const double ratio = cfg.safety * gi * gp;
  const double sf = fmax(cfg.factor_min, fmin(ratio, cfg.factor_max));
  if (ep >= 1.0) cstate = ep;
  return sf * dt;
}

// ------------------------------------------------------------ slot access
struct SlotV {
  double *sc, *m, *L, *cA, *cm, *cL, *ctl, *cto, *b, *J;
};
template <int G>
__device__ __forceinline__ SlotV slotv(const Ctx<G>& c, int s) {
  double* base = c.slot(s);
  const Layout& L = *c.L;
  return SlotV{base + L.s_sc, base + L.s_m, base + L.s_L, base + L.s_cA, base + L.s_cm, base + L.s_cL,
               base + L.s_ctl, base + L.s_cto, base + L.s_b, base + L.s_J};
}

template <int G>
__device__ void copy_slot(const Ctx<G>& c, int dst, int src) {
  g_copy(c.g, c.slot(dst), c.slot(src), c.L->s_size);
  c.g.sync();
}

// ------------------------------------------------------------- the step
// solver_dynamic.step (solvers.py:549-599), solver.step (:702-733) and
// solver_mle.step (:395-437), followed by the residual error estimate.
// Returns error_power (only meaningful with `estimate`).
template <int G>
__device__ double step(const Ctx<G>& c, int from, int to, double dt, bool estimate) {
  const Layout& L = *c.L;
  const pdq_config& cfg = c.a->cfg;
  SlotV S = slotv(c, from), R = slotv(c, to);
  double* P = c.W(L.w_P);
  double* PI = c.W(L.w_PI);
  double* big = c.W(L.w_big);
  double* Um = c.W(L.w_Um);
  double* UL = c.W(L.w_UL);
  double* Pm = c.W(L.w_Pm);
  double* PL = c.W(L.w_PL);
  double* om = c.W(L.w_om);
  double* oL = c.W(L.w_oL);
  double* gain = c.W(L.w_G);
  double* mo = c.W(L.w_mo);
  double* tmp = c.W(L.w_T1);
  double* y0buf = c.W(L.w_y0);
  const double t = S.sc[SC_T];
  const double tn = t + dt;
  const bool smoother = cfg.strategy != PDQ_FILTER;
  const bool ident = S.sc[SC_IDENT] != 0.0;
  const bool state_std = cfg.error_estimator == PDQ_ERROR_STATE_STD;
  precon(c, dt, P, PI);
  c.g.sync();
  double sigma = 1.0, ep = 0.0;
  if (cfg.solver == PDQ_SOLVER_DYNAMIC) {
    const double s1 = noise_scale(c, dt, 1.0);
    apply_flat_mean(c, P, PI, S.m, Um);
    apply_flat_chol(c, P, s1, UL);
    c.g.sync();
    linearize(c, Um, tn, R.b, R.J, y0buf);
    marginalise_fx(c, R.b, R.J, Um, UL, om, oL, big);
    sigma = rms_zero(c, om, oL);
    // error_state_std from the step's (first) linearisation, before the
    // prediction takes Pm / PL; the errors stay in om
    if (estimate && state_std) state_std_errors(c, R.b, R.J, Um, UL, Pm, PL, big, gain, mo, oL, om);
    const double s = noise_scale(c, dt, sigma);
    if (!smoother) {
      marginalise_prior(c, P, PI, s, S.m, S.L, Pm, PL, big);
    } else {
      double* bA = c.W(L.w_bA);
      double* bm = c.W(L.w_bm);
      double* bL = c.W(L.w_bL);
      double* btl = c.W(L.w_btl);
      double* bto = c.W(L.w_bto);
      revert_prior(c, P, PI, s, S.m, S.L, Pm, PL, bA, bm, bL, btl, bto, big, tmp);
      c.g.sync();
      if (cfg.strategy == PDQ_FIXEDPOINT) {
        merge(c, ident, S.cA, S.cm, S.cL, S.ctl, S.cto, bA, bm, bL, btl, bto, R.cA, R.cm, R.cL, R.ctl, R.cto, big);
      } else {
        g_copy(c.g, R.cA, bA, c.nc * c.nc);
        g_copy(c.g, R.cm, bm, c.nc * c.mc);
        g_copy(c.g, R.cL, bL, c.nc * c.nc);
        g_copy(c.g, R.ctl, btl, c.nc);
        g_copy(c.g, R.cto, bto, c.nc);
      }
    }
    c.g.sync();
    // solvers.py:576-582: linearise again at the calibrated prediction
    if (cfg.re_linearize_after_calibration) linearize(c, Pm, tn, R.b, R.J, y0buf, 1);
    update(c, R.b, R.J, Pm, PL, R.m, R.L, big, gain, mo, nullptr);
    if (estimate) ep = state_std ? error_power_state(c, om, S.m, R.m, dt) : error_power(c, oL, sigma, S.m, R.m, dt);
  } else {
    const double s1 = noise_scale(c, dt, 1.0);
    if (!smoother) {
      marginalise_prior(c, P, PI, s1, S.m, S.L, Pm, PL, big);
    } else {
      double* bA = c.W(L.w_bA);
      double* bm = c.W(L.w_bm);
      double* bL = c.W(L.w_bL);
      double* btl = c.W(L.w_btl);
      double* bto = c.W(L.w_bto);
      revert_prior(c, P, PI, s1, S.m, S.L, Pm, PL, bA, bm, bL, btl, bto, big, tmp);
      c.g.sync();
      if (cfg.strategy == PDQ_FIXEDPOINT) {
        merge(c, ident, S.cA, S.cm, S.cL, S.ctl, S.cto, bA, bm, bL, btl, bto, R.cA, R.cm, R.cL, R.ctl, R.cto, big);
      } else {
        g_copy(c.g, R.cA, bA, c.nc * c.nc);
        g_copy(c.g, R.cm, bm, c.nc * c.mc);
        g_copy(c.g, R.cL, bL, c.nc * c.nc);
        g_copy(c.g, R.ctl, btl, c.nc);
        g_copy(c.g, R.cto, bto, c.nc);
      }
    }
    c.g.sync();
    linearize(c, Pm, tn, R.b, R.J, y0buf);
    const bool mle = cfg.solver == PDQ_SOLVER_MLE;
    update(c, R.b, R.J, Pm, PL, R.m, R.L, big, gain, mo, mle ? oL : nullptr);
    if (mle) {
      // bayes_rule_and_residual_whitened_rms_tree: observed = (mo, R_Y^T)
      const double nt = rms_zero(c, mo, oL);
      const double nd = S.sc[SC_MLE_N];
      const double x1 = sqrt(nd / (nd + 1)) * S.sc[SC_MLE_RUN];
      const double x2 = sqrt(1 / (nd + 1)) * nt;
      c.g.sync();
      if (c.g.l == 0) {
        R.sc[SC_MLE_RUN] = hypot(x1, x2);
        R.sc[SC_MLE_N] = nd + 1;
      }
    }
    if (estimate) {
      apply_flat_mean(c, P, PI, S.m, Um);
      apply_flat_chol(c, P, s1, UL);
      c.g.sync();
      // re_linearize_before_error (solvers.py:951-956): linearise at the
      // extrapolation instead of reusing the step's linearisation
      if (cfg.error_re_linearize) linearize(c, Um, tn, R.b, R.J, y0buf);
      if (state_std) {
        state_std_errors(c, R.b, R.J, Um, UL, Pm, PL, big, gain, mo, oL, om);
        ep = error_power_state(c, om, S.m, R.m, dt);
      } else {
        marginalise_fx(c, R.b, R.J, Um, UL, om, oL, big);
        const double se = rms_zero(c, om, oL);
        ep = error_power(c, oL, se, S.m, R.m, dt);
      }
    }
    sigma = mle ? S.sc[SC_OS] : 1.0;
  }
  c.g.sync();
  if (c.g.l == 0) {
    R.sc[SC_T] = tn;
    R.sc[SC_OS] = sigma;
    R.sc[SC_NSTEPS] = S.sc[SC_NSTEPS] + 1.0;
    R.sc[SC_IDENT] = 0.0;
    if (cfg.solver != PDQ_SOLVER_MLE) {
      R.sc[SC_MLE_RUN] = 0.0;
      R.sc[SC_MLE_N] = 0.0;
    }
  }
  c.g.sync();
  return ep;
}

// ---------------------------------------------------------------- outputs
template <int G>
__device__ void write_normal(const Ctx<G>& c, const double* m, const double* Lc, double* gm, double* gL) {
  if (gm) {
    const int mcp = c.mcp, mc = c.mc;
    for (int e = c.g.l; e < c.nc * mcp; e += G) {
      const int i = e / mcp, col = e % mcp;
      gm[e] = m[i * mc + col];
    }
  }
  if (gL) g_copy(c.g, gL, Lc, c.nc * c.nc);
}

template <int G>
__device__ void write_cond(const Ctx<G>& c, const SlotV& S, int inst, int j) {
  const pdq_solution& o = c.a->out;
  const int nc = c.nc, Km1 = c.a->K - 1, ms = c.nc * c.mcp;
  const size_t idx = (size_t)inst * Km1 + j;
  const bool ident = S.sc[SC_IDENT] != 0.0;
  for (int e = c.g.l; e < nc * nc; e += G) {
    const int i = e / nc, k = e % nc;
    o.cond_A[idx * nc * nc + e] = ident ? (i == k ? 1.0 : 0.0) : S.cA[e];
    o.cond_chol[idx * nc * nc + e] = ident ? 0.0 : S.cL[e];
  }
  for (int e = c.g.l; e < ms; e += G) {
    const int i = e / c.mcp, col = e % c.mcp;
    o.cond_mean[idx * ms + e] = ident ? 0.0 : S.cm[i * c.mc + col];
  }
  for (int i = c.g.l; i < nc; i += G) {
    o.cond_tl[idx * nc + i] = ident ? 1.0 : S.ctl[i];
    o.cond_to[idx * nc + i] = ident ? 1.0 : S.cto[i];
  }
}

// Writes checkpoint k of instance `inst` from a solution view; `u` is the
// marginal written as u (filter) or as the filtering marginal (smoothers).
template <int G>
__device__ void write_checkpoint(const Ctx<G>& c, int inst, int k, double t, const double* m, const double* Lc,
                                 double os, double nsteps, const SlotV* condsrc) {
  const pdq_solution& o = c.a->out;
  const int K = c.a->K, ms = c.nc * c.mcp, cs = c.nc * c.nc;
  const bool smoother = c.a->cfg.strategy != PDQ_FILTER;
  const size_t row = (size_t)inst * K + k;
  if (c.g.l == 0) {
    if (o.t) o.t[row] = t;
    const bool dyn = c.a->cfg.solver == PDQ_SOLVER_DYNAMIC;
    if (o.output_scale && dyn) o.output_scale[row] = os;
    if (o.num_steps && k > 0) o.num_steps[(size_t)inst * (K - 1) + k - 1] = (int64_t)nsteps;
  }
  if (smoother) {
    write_normal(c, m, Lc, o.filt_mean ? o.filt_mean + row * ms : nullptr, o.filt_chol ? o.filt_chol + row * cs : nullptr);
    if (k > 0 && condsrc) write_cond(c, *condsrc, inst, k - 1);
  } else {
    write_normal(c, m, Lc, o.mean ? o.mean + row * ms : nullptr, o.chol ? o.chol + row * cs : nullptr);
  }
}

// ------------------------------------------------------------ finalize
// userfriendly_output + strategy.finalize (solvers.py:439-480, 601-633,
// 735-767; estimators_and_losses.py:380-391, 437-470).
template <int G>
__device__ void finalize(const Ctx<G>& c, int inst, int sf) {
  const Layout& L = *c.L;
  const pdq_config& cfg = c.a->cfg;
  const pdq_solution& o = c.a->out;
  const int K = c.a->K, nc = c.nc, ms = nc * c.mcp, cs = nc * nc, Km1 = K - 1;
  SlotV S = slotv(c, sf);
  __threadfence_block();
  c.g.sync();
  double scale = 1.0;
  const bool mle = cfg.solver == PDQ_SOLVER_MLE;
  if (mle) {
    scale = S.sc[SC_MLE_RUN];
    if (cfg.correct_asymptotic_underconfidence) {
      const double ns = (double)o.num_steps[(size_t)inst * Km1 + Km1 - 1];
      scale = scale / sqrt(ns);
    }
    if (o.output_scale && c.g.l == 0)
      for (int k = 0; k < Km1; ++k) o.output_scale[(size_t)inst * K + k] = scale;
  } else if (cfg.solver == PDQ_SOLVER && o.output_scale && c.g.l == 0) {
    for (int k = 0; k < Km1; ++k) o.output_scale[(size_t)inst * K + k] = 1.0;
  }
  if (cfg.strategy == PDQ_FILTER) {
    if (mle && o.chol)
      for (int e = c.g.l; e < K * cs; e += G) o.chol[(size_t)inst * K * cs + e] = scale * o.chol[(size_t)inst * K * cs + e];
    return;
  }
  if (mle) {
    for (int e = c.g.l; e < K * cs; e += G) o.filt_chol[(size_t)inst * K * cs + e] = scale * o.filt_chol[(size_t)inst * K * cs + e];
    for (int e = c.g.l; e < Km1 * cs; e += G)
      o.cond_chol[(size_t)inst * Km1 * cs + e] = scale * o.cond_chol[(size_t)inst * Km1 * cs + e];
    for (int e = c.g.l; e < nc * nc; e += G) {
      S.L[e] = scale * S.L[e];
      S.cL[e] = scale * S.cL[e];
    }
    __threadfence_block();
    c.g.sync();
  }
  double* big = c.W(L.w_big);
  double* Xm = c.W(L.w_Pm);
  double* XL = c.W(L.w_PL);
  double* Ym = c.W(L.w_Um);
  double* YL = c.W(L.w_UL);
  // rv_at_t1 = posterior1.conditional.marginalise(posterior1.marginal)
  if (S.sc[SC_IDENT] != 0.0) {
    g_copy(c.g, Xm, S.m, nc * c.mc);
    g_copy(c.g, XL, S.L, nc * nc);
  } else {
    marginalise_cond(c, S.cA, S.cm, S.cL, S.ctl, S.cto, S.m, S.L, Xm, XL, big);
  }
  c.g.sync();
  const size_t rowK = (size_t)inst * K + (K - 1);
  write_normal(c, Xm, XL, o.term_mean ? o.term_mean + (size_t)inst * ms : nullptr,
               o.term_chol ? o.term_chol + (size_t)inst * cs : nullptr);
  write_normal(c, Xm, XL, o.mean ? o.mean + rowK * ms : nullptr, o.chol ? o.chol + rowK * cs : nullptr);
  // backward pass: x_j = cond_j.marginalise(x_{j+1}), j = K-2 .. 0
  double* cA = c.W(L.w_bA);
  double* cm = c.W(L.w_bm);
  double* cL = c.W(L.w_bL);
  double* ctl = c.W(L.w_btl);
  double* cto = c.W(L.w_bto);
  for (int j = Km1 - 1; j >= 0; --j) {
    const size_t idx = (size_t)inst * Km1 + j;
    __threadfence_block();
    g_copy(c.g, cA, o.cond_A + idx * nc * nc, nc * nc);
    g_copy(c.g, cL, o.cond_chol + idx * nc * nc, nc * nc);
    for (int e = c.g.l; e < nc * c.mc; e += G) {
      const int i = e / c.mc, col = e % c.mc;
      cm[e] = col < c.mcp ? o.cond_mean[idx * ms + i * c.mcp + col] : 0.0;
    }
    g_copy(c.g, ctl, o.cond_tl + idx * nc, nc);
    g_copy(c.g, cto, o.cond_to + idx * nc, nc);
    c.g.sync();
    marginalise_cond(c, cA, cm, cL, ctl, cto, Xm, XL, Ym, YL, big);
    c.g.sync();
    const size_t row = (size_t)inst * K + j;
    write_normal(c, Ym, YL, o.mean ? o.mean + row * ms : nullptr, o.chol ? o.chol + row * cs : nullptr);
    g_copy(c.g, Xm, Ym, nc * c.mc);
    g_copy(c.g, XL, YL, nc * nc);
    c.g.sync();
  }
}

// ------------------------------------------------------ interpolations
// ProbabilisticSolver.interpolate_fwd (solvers.py:205-269) for the three
// strategies.  from = interp_from slot, to = step_from slot; writes the
// checkpoint output and updates both slots.
template <int G>
__device__ void interp_beyond(const Ctx<G>& c, int inst, int k, double t1, int from, int to) {
  const Layout& L = *c.L;
  const pdq_config& cfg = c.a->cfg;
  SlotV F = slotv(c, from), T = slotv(c, to);
  double* P = c.W(L.w_P);
  double* PI = c.W(L.w_PI);
  double* big = c.W(L.w_big);
  double* Pm = c.W(L.w_Pm);
  double* PL = c.W(L.w_PL);
  double* tmp = c.W(L.w_T1);
  const double os = T.sc[SC_OS];
  const double dt0 = t1 - F.sc[SC_T];
  const double dt1 = T.sc[SC_T] - t1;
  precon(c, dt0, P, PI);
  c.g.sync();
  const double s0 = noise_scale(c, dt0, os);
  if (cfg.strategy == PDQ_FILTER) {
    marginalise_prior(c, P, PI, s0, F.m, F.L, Pm, PL, big);
    c.g.sync();
    write_checkpoint(c, inst, k, t1, Pm, PL, os, T.sc[SC_NSTEPS], nullptr);
    g_copy(c.g, F.m, Pm, c.nc * c.mc);
    g_copy(c.g, F.L, PL, c.nc * c.nc);
    if (c.g.l == 0) F.sc[SC_T] = t1;
    c.g.sync();
    return;
  }
  double* bA = c.W(L.w_bA);
  double* bm = c.W(L.w_bm);
  double* bL = c.W(L.w_bL);
  double* btl = c.W(L.w_btl);
  double* bto = c.W(L.w_bto);
  double* eA = c.W(L.w_eA);
  double* em = c.W(L.w_em);
  double* eL = c.W(L.w_eL);
  double* etl = c.W(L.w_etl);
  double* eto = c.W(L.w_eto);
  // extrapolate t0 -> t: ext_t (marginal Pm/PL, conditional e*)
  revert_prior(c, P, PI, s0, F.m, F.L, Pm, PL, bA, bm, bL, btl, bto, big, tmp);
  c.g.sync();
  if (cfg.strategy == PDQ_FIXEDPOINT) {
    merge(c, F.sc[SC_IDENT] != 0.0, F.cA, F.cm, F.cL, F.ctl, F.cto, bA, bm, bL, btl, bto, eA, em, eL, etl, eto, big);
  } else {
    g_copy(c.g, eA, bA, c.nc * c.nc);
    g_copy(c.g, em, bm, c.nc * c.mc);
    g_copy(c.g, eL, bL, c.nc * c.nc);
    g_copy(c.g, etl, btl, c.nc);
    g_copy(c.g, eto, bto, c.nc);
  }
  c.g.sync();
  // extrapolate t -> t1 from (ext_t.marginal, identity) (fixed point) or
  // from ext_t (fixed interval); only the backward conditional is kept.
  precon(c, dt1, P, PI);
  c.g.sync();
  const double s1 = noise_scale(c, dt1, os);
  double* Ym = c.W(L.w_Um);
  double* YL = c.W(L.w_UL);
  revert_prior(c, P, PI, s1, Pm, PL, Ym, YL, bA, bm, bL, btl, bto, big, tmp);
  c.g.sync();
  // new step_from.cond = (fixed point) identity o bw, (fixed interval) bw
  if (cfg.strategy == PDQ_FIXEDPOINT) {
    merge(c, true, nullptr, nullptr, nullptr, nullptr, nullptr, bA, bm, bL, btl, bto, T.cA, T.cm, T.cL, T.ctl, T.cto, big);
  } else {
    g_copy(c.g, T.cA, bA, c.nc * c.nc);
    g_copy(c.g, T.cm, bm, c.nc * c.mc);
    g_copy(c.g, T.cL, bL, c.nc * c.nc);
    g_copy(c.g, T.ctl, btl, c.nc);
    g_copy(c.g, T.cto, bto, c.nc);
  }
  // output: interpolated = (ext_t.marginal, ext_t.conditional)
  SlotV E{nullptr, nullptr, nullptr, eA, em, eL, etl, eto, nullptr, nullptr};
  double esc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  E.sc = esc;
  c.g.sync();
  write_checkpoint(c, inst, k, t1, Pm, PL, os, T.sc[SC_NSTEPS], &E);
  // new interp_from: fixed point (ext_t.marginal, identity); fixed interval ext_t
  g_copy(c.g, F.m, Pm, c.nc * c.mc);
  g_copy(c.g, F.L, PL, c.nc * c.nc);
  if (cfg.strategy == PDQ_FIXEDINTERVAL) {
    g_copy(c.g, F.cA, eA, c.nc * c.nc);
    g_copy(c.g, F.cm, em, c.nc * c.mc);
    g_copy(c.g, F.cL, eL, c.nc * c.nc);
    g_copy(c.g, F.ctl, etl, c.nc);
    g_copy(c.g, F.cto, eto, c.nc);
  }
  c.g.sync();
  if (c.g.l == 0) {
    F.sc[SC_T] = t1;
    F.sc[SC_IDENT] = cfg.strategy == PDQ_FIXEDPOINT ? 1.0 : 0.0;
    T.sc[SC_IDENT] = 0.0;
  }
  c.g.sync();
}

// ProbabilisticSolver.interpolate_fwd_at_t1 (solvers.py:271-315).
template <int G>
__device__ void interp_at(const Ctx<G>& c, int inst, int k, int from, int to) {
  const pdq_config& cfg = c.a->cfg;
  SlotV T = slotv(c, to);
  write_checkpoint(c, inst, k, T.sc[SC_T], T.m, T.L, T.sc[SC_OS], T.sc[SC_NSTEPS], &T);
  c.g.sync();
  // new interp_from: posterior_t1 (filter / fixed interval) or the reset one
  copy_slot(c, from, to);
  if (cfg.strategy == PDQ_FIXEDPOINT) {
    SlotV F = slotv(c, from);
    if (c.g.l == 0) {
      T.sc[SC_IDENT] = 1.0;
      F.sc[SC_IDENT] = 1.0;
    }
  }
  c.g.sync();
}


// ------------------------------------------------------------ LML + grad
// Marginal log-likelihood of loss_lml_timeseries (estimators_and_losses.py:
// 53-105) on the fixed-point smoother's Markov sequence: forward pass on a
// fixed grid with a checkpoint (interpolate_fwd_at_t1 reset, :534-547) every
// `every` grid points, then MarkovSequence.evaluate_lml's reverse scan
// (:180-218) with lstsq gains.  Forward-mode tangents (d/d params, d/d y0)
// ride along as extra mean columns: with ts0 the covariance recursion does
// not depend on them, so every square-root factor is shared and only the
// mean recursions and the log-densities carry derivatives.
template <int G>
__device__ void lml_instance(const Ctx<G>& c, int inst) {
  const pdq_config& cfg = c.a->cfg;
  const KArgs& a = *c.a;
  const Layout& L = *c.L;
  const int N = c.N, d = c.d, nc = c.nc, mc = c.mc, mcp = c.mcp, ko = c.ko;
  const int ntan = mc / mcp - 1;
  const int np = cfg.n_params;
  const int K = a.K, every = a.every, Kobs = (K - 1) / every + 1;
  for (int i = c.g.l; i < np; i += G) c.prm()[i] = a.params[(size_t)inst * np + i];
  int sf = 0, pr = 2;
  g_fill(c.g, c.slot(sf), 0.0, L.s_size);
  c.g.sync();
  SlotV S0 = slotv(c, sf);
  // prior mean: Taylor coefficients of y0 with their tangents, computed by
  // taylor_tangent_kernel (layout n_inst x (N d) x (1 + ntan))
  const double* tc = a.tcoeffs + (size_t)inst * N * d * (1 + ntan);
  for (int e = c.g.l; e < N * d * (1 + ntan); e += G) {
    const int ik = e / (1 + ntan), q = e % (1 + ntan);
    const int i = ik / d, k = ik % d;
    if (c.dense) S0.m[(i * d + k) * mc + q] = tc[e];
    else S0.m[i * mc + q * d + k] = tc[e];
  }
  for (int i = c.g.l; i < nc; i += G) S0.L[i * nc + i] = cfg.init_std_rows ? cfg.init_std_rows[i] : cfg.init_std;
  if (c.g.l == 0) {
    S0.sc[SC_T] = a.times[0];
    S0.sc[SC_OS] = 1.0;
    S0.sc[SC_IDENT] = 1.0;
  }
  c.g.sync();
  const int csz = 2 * nc * nc + nc * mc + 2 * nc;
  double* store = a.scratch + (size_t)inst * (Kobs - 1) * csz;
  for (int k = 1; k < K; ++k) {
    c.lin_set = pdq_lin_set(c.a->cfg, k - 1);
    step(c, sf, pr, a.times[k] - a.times[k - 1], false);
    const int t = sf;
    sf = pr;
    pr = t;
    if (k % every == 0) {
      SlotV S = slotv(c, sf);
      double* dst = store + (size_t)(k / every - 1) * csz;
      g_copy(c.g, dst, S.cA, nc * nc);
      g_copy(c.g, dst + nc * nc, S.cm, nc * mc);
      g_copy(c.g, dst + nc * nc + nc * mc, S.cL, nc * nc);
      g_copy(c.g, dst + 2 * nc * nc + nc * mc, S.ctl, nc);
      g_copy(c.g, dst + 2 * nc * nc + nc * mc + nc, S.cto, nc);
      c.g.sync();
      if (c.g.l == 0) S.sc[SC_IDENT] = 1.0;
      c.g.sync();
    }
  }
  __threadfence_block();
  if (cfg.solver == PDQ_SOLVER_MLE) {
    // solver_mle's userfriendly_output rescales the posterior by the
    // calibrated output scale before the loss sees it (solvers.py:439-480,
    // strategy.finalize): every stored conditional's noise factor and the
    // terminal marginal and conditional.  With tangents refused, the factors
    // are the only thing the scale touches.
    c.g.sync();
    SlotV S = slotv(c, sf);
    double scale = S.sc[SC_MLE_RUN];
    if (cfg.correct_asymptotic_underconfidence) scale = scale / sqrt((double)(K - 1));
    for (int e = c.g.l; e < nc * nc; e += G) {
      S.L[e] = scale * S.L[e];
      S.cL[e] = scale * S.cL[e];
    }
    for (int j = 0; j < Kobs - 1; ++j) {
      double* nl = store + (size_t)j * csz + nc * nc + nc * mc;
      for (int e = c.g.l; e < nc * nc; e += G) nl[e] = scale * nl[e];
    }
    __threadfence_block();
    c.g.sync();
  }
  // terminal marginal: posterior1.conditional.marginalise(posterior1.marginal)
  double* big = c.W(L.w_big);
  double* Xm = c.W(L.w_Pm);
  double* XL = c.W(L.w_PL);
  double* Ym = c.W(L.w_Um);
  double* YL = c.W(L.w_UL);
  SlotV S = slotv(c, sf);
  if (S.sc[SC_IDENT] != 0.0) {
    g_copy(c.g, Xm, S.m, nc * mc);
    g_copy(c.g, XL, S.L, nc * nc);
  } else {
    marginalise_cond(c, S.cA, S.cm, S.cL, S.ctl, S.cto, S.m, S.L, Xm, XL, big);
  }
  c.g.sync();
  double* gain = c.W(L.w_G);
  double* mo = c.W(L.w_mo);
  double* lp = c.W(L.w_lml);          // 1 + ntan running log-densities
  double* lpb = lp + (1 + ntan);       // this node's terms
  double* w = lpb + (1 + ntan);        // ko whitened residuals
  double* wt = w + ko;                 // ko tangents of them
  if (c.g.l == 0)
    for (int q = 0; q <= ntan; ++q) lp[q] = 0.0;
  double num = 0.0;
  const double c2pi = log(M_PI * 2.0);
  for (int j = Kobs - 1; j >= 0; --j) {
    if (j < Kobs - 1) {
      const double* src = store + (size_t)j * csz;
      marginalise_cond(c, src, src + nc * nc, src + nc * nc + nc * mc, src + 2 * nc * nc + nc * mc,
                       src + 2 * nc * nc + nc * mc + nc, Xm, XL, Ym, YL, big);
      c.g.sync();
      g_copy(c.g, Xm, Ym, nc * mc);
      g_copy(c.g, XL, YL, nc * nc);
      c.g.sync();
    }
    // model_j = marginal.to_derivative(i, std_j) (ssm_impl_dense.py:206-217,
    // ssm_impl_isotropic.py:271-278) with i = lml_tcoeff_index: the observed
    // rows are i d + k (dense) and i (isotropic), i.e. ro + o with ro = i ko;
    // bayes_rule_and_logpdf_tree
    const double* y = a.data + ((size_t)inst * Kobs + j) * d;
    const double* sd = a.stds + ((size_t)inst * Kobs + j) * a.n_std;
    const int n2 = ko + nc, ro = cfg.lml_tcoeff_index * ko;
    for (int e = c.g.l; e < n2 * n2; e += G) {
      const int r = e / n2, col = e % n2;
      double v;
      if (r < ko) v = col < ko ? (r == col ? sd[a.n_std == 1 ? 0 : r] : 0.0) : 0.0;
      else v = col < ko ? XL[(ro + col) * nc + (r - ko)] : XL[(col - ko) * nc + (r - ko)];
      big[e] = v;
    }
    for (int e = c.g.l; e < ko * mc; e += G) {
      const int o = e / mc, col = e % mc;
      mo[e] = Xm[(ro + o) * mc + col] + 0.0;
    }
    c.g.sync();
    g_qr_r(c.g, big, n2, n2, n2);
    g_gain(c.g, big, n2, ko, nc, gain, cfg.lml_solve == PDQ_LML_SOLVE_TRIU ? 0 : 1, c.W(L.w_svd));
    c.g.sync();
    if (c.g.l == 0) {
      // observed.logpdf (ssm_impl_dense.py:170-180 / isotropic :217-236);
      // qr_r of the already-triangular R_Y leaves it unchanged.
      for (int q = 0; q <= ntan; ++q) lpb[q] = 0.0;
      double slogdet = 0.0;
      for (int i = 0; i < ko; ++i) slogdet += log(fabs(big[i * n2 + i]));
      for (int col = 0; col < mcp; ++col) {
        for (int i = 0; i < ko; ++i) {
          const double yi = c.dense ? y[i] : y[col];
          double s = yi - mo[i * mc + col];
          for (int k = 0; k < i; ++k) s -= big[k * n2 + i] * w[k];
          w[i] = s / big[i * n2 + i];
        }
        double sq = 0.0;
        for (int i = 0; i < ko; ++i) sq += w[i] * w[i];
        if (c.dense) lpb[0] += -1.0 / 2.0 * sq - (double)ko / 2.0 * c2pi - slogdet;
        else lpb[0] += -0.5 * (2.0 * slogdet + sq + (double)ko * c2pi);
        for (int q = 1; q <= ntan; ++q) {
          double dot = 0.0;
          for (int i = 0; i < ko; ++i) {
            double s = -mo[i * mc + q * mcp + col];
            for (int k = 0; k < i; ++k) s -= big[k * n2 + i] * wt[k];
            wt[i] = s / big[i * n2 + i];
            dot += w[i] * wt[i];
          }
          lpb[q] += -dot;
        }
      }
      for (int q = 0; q <= ntan; ++q) {
        const double v = lpb[q];
        if (num == 0.0) lp[q] = v;
        else if (a.average_pdfs) lp[q] = (lp[q] * num + v) / (num + 1);
        else lp[q] = lp[q] + v;
      }
    }
    c.g.sync();
    num += 1.0;
    // updated = reverted.apply_flat(data)
    for (int e = c.g.l; e < nc * mc; e += G) {
      const int i = e / mc, col = e % mc;
      double gy = 0.0, gm = 0.0;
      for (int o = 0; o < ko; ++o) {
        const double yo = col < mcp ? (c.dense ? y[o] : y[col]) : 0.0;
        gy += gain[i * ko + o] * yo;
        gm += gain[i * ko + o] * mo[o * mc + col];
      }
      Ym[e] = c.dense ? 1.0 * (gy + (Xm[e] - gm)) : gy + (Xm[e] - gm);
    }
    for (int e = c.g.l; e < nc * nc; e += G) {
      const int i = e / nc, k = e % nc;
      YL[e] = big[(ko + k) * n2 + ko + i];
    }
    c.g.sync();
    g_copy(c.g, Xm, Ym, nc * mc);
    g_copy(c.g, XL, YL, nc * nc);
    c.g.sync();
  }
  if (c.g.l == 0) {
    if (a.lml_value) a.lml_value[inst] = lp[0];
    if (a.lml_grad)
      for (int q = 0; q < ntan; ++q) a.lml_grad[(size_t)inst * ntan + q] = lp[1 + q];
  }
}

// --------------------------------------------------------------- kernel
template <int G>
__device__ void init_instance(const Ctx<G>& c, int inst, int slot0) {
  const pdq_config& cfg = c.a->cfg;
  SlotV S = slotv(c, slot0);
  const int N = c.N, d = c.d, nc = c.nc, mc = c.mc;
  g_fill(c.g, c.slot(slot0), 0.0, c.L->s_size);
  c.g.sync();
  // prior init: mean = tcoeffs, chol = init_std * I (from_mean_and_std)
  const double* tc = c.a->tcoeffs + (size_t)inst * N * d;
  for (int e = c.g.l; e < N * d; e += G) {
    const int i = e / d, k = e % d;
    if (c.dense) S.m[(i * d + k) * mc] = tc[e];
    else S.m[i * mc + k] = tc[e];
  }
  for (int i = c.g.l; i < nc; i += G) S.L[i * nc + i] = cfg.init_std_rows ? cfg.init_std_rows[i] : cfg.init_std;
  if (c.g.l == 0) {
    S.sc[SC_T] = c.a->times[(size_t)inst * c.a->K];
    S.sc[SC_OS] = 1.0;
    S.sc[SC_NSTEPS] = 0.0;
    S.sc[SC_IDENT] = cfg.strategy != PDQ_FILTER ? 1.0 : 0.0;
  }
  c.g.sync();
}

template <int G>
__device__ void solve_instance(const Ctx<G>& c, int inst) {
  const pdq_config& cfg = c.a->cfg;
  const pdq_solution& o = c.a->out;
  const int K = c.a->K;
  const double* ts = c.a->times + (size_t)inst * K;
  // vector-field parameters
  for (int i = c.g.l; i < cfg.n_params; i += G) c.prm()[i] = c.a->params[(size_t)inst * cfg.n_params + i];
  int sf = 0, ifr = 1, pr = 2;
  init_instance(c, inst, sf);
  copy_slot(c, ifr, sf);
  {
    SlotV S = slotv(c, sf);
    write_checkpoint(c, inst, 0, S.sc[SC_T], S.m, S.L, 1.0, 0.0, nullptr);
  }
  long attempts = 0, accepted = 0;
  int status = PDQ_STATUS_OK;
  if (c.a->mode == MODE_FIXED_GRID) {
    for (int k = 1; k < K && status == PDQ_STATUS_OK; ++k) {
      const double dt = ts[k] - ts[k - 1];
      c.lin_set = pdq_lin_set(c.a->cfg, k - 1);
      step(c, sf, pr, dt, false);
      ++attempts;
      ++accepted;
      SlotV R = slotv(c, pr);
      write_checkpoint(c, inst, k, R.sc[SC_T], R.m, R.L, R.sc[SC_OS], R.sc[SC_NSTEPS], &R);
      c.g.sync();
      const int tmp = sf;
      sf = pr;
      pr = tmp;
    }
  } else {
    double dt = cfg.dt0_per_instance ? cfg.dt0_per_instance[inst] : cfg.dt0, cstate = 1.0;
    for (int k = 1; k < K && status == PDQ_STATUS_OK; ++k) {
      const double t1 = ts[k];
      bool cont = true;
      while (cont) {
        if (slotv(c, sf).sc[SC_T] + cfg.eps < t1) {
          double ep = 0.9;
          int att = 0;
          while (ep < 1.0) {
            double dtc = dt;
            if (cfg.clip_dt) dtc = fmin(dtc, t1 - slotv(c, sf).sc[SC_T]);
            c.lin_set = pdq_lin_set(c.a->cfg, accepted);
            ep = step(c, sf, pr, dtc, true);
            dt = control_apply(cfg, dtc, cstate, ep);
            ++att;
            ++attempts;
            if (!isfinite(ep) || !isfinite(dt)) {
              status = PDQ_STATUS_NONFINITE;
              break;
            }
            if (ep < 1.0 && (dt == 0.0 || slotv(c, sf).sc[SC_T] + dt == slotv(c, sf).sc[SC_T])) {
              status = PDQ_STATUS_DT_UNDERFLOW;
              break;
            }
            if (ep >= 1.0 && slotv(c, pr).sc[SC_T] == slotv(c, sf).sc[SC_T]) {
              status = PDQ_STATUS_DT_UNDERFLOW;  // accepted without advancing t
              break;
            }
            if ((cfg.max_attempts > 0 && att >= cfg.max_attempts && ep < 1.0) || attempts >= kMaxTotalAttempts) {
              status = PDQ_STATUS_MAX_ATTEMPTS;
              break;
            }
          }
          if (status != PDQ_STATUS_OK) break;
          ++accepted;
          // interp_from <- step_from, step_from <- proposed
          const int old_if = ifr;
          ifr = sf;
          sf = pr;
          pr = old_if;
        }
        const double tsf = slotv(c, sf).sc[SC_T];
        if (tsf + cfg.eps < t1) {
          cont = true;  // interp_skip
        } else if (tsf > t1 + cfg.eps) {
          interp_beyond(c, inst, k, t1, ifr, sf);
          cont = slotv(c, sf).sc[SC_T] + cfg.eps < t1;
        } else {
          interp_at(c, inst, k, ifr, sf);
          cont = false;
        }
      }
    }
  }
  if (status == PDQ_STATUS_OK) finalize(c, inst, sf);
  if (c.g.l == 0) {
    if (o.attempts) o.attempts[inst] = attempts;
    if (o.accepted) o.accepted[inst] = accepted;
    if (o.status) o.status[inst] = status;
  }
}

// ---------------------------------------------------------- off-grid marginals
// ProbabilisticSolver.offgrid_marginals (solvers.py:149-203) for the
// factorisations without a lane kernel (dense; isotropic of other shapes),
// one G-lane group per query time, with the solve's primitives and slot
// workspace.  For t the group finds the grid interval (t0, t1] by
// searchsorted(t_grid, t) (side "left", as the reference; the index is
// clamped to [1, K - 1], so every access stays inside the arrays) and
// evaluates
// * the filter (estimators_and_losses.py:403-414): the marginal at t0
//   predicted to t with the transition of t1's output scale;
// * the fixed-interval smoother (:674-707): the filtering marginal at t0
//   reverted to t and on to t1, then the smoothed marginal at t1
//   marginalised back through the conditional t1 -> t.
template <int G>
__device__ void offgrid_query(const Ctx<G>& c, const OffgridCoopArgs& o, size_t q) {
  const Layout& L = *c.L;
  const int K = c.a->K, nc = c.nc, mc = c.mc, ms = nc * mc, cs = nc * nc;
  const size_t inst = q / (size_t)o.n_query;
  const double* ts = c.a->times + inst * K;
  const double t = o.t_query[q];
  int lo = 0, hi = K;
  while (lo < hi) {  // first index with ts[i] >= t
    const int mid = (lo + hi) / 2;
    if (ts[mid] < t) lo = mid + 1;
    else hi = mid;
  }
  const int i = lo < 1 ? 1 : (lo > K - 1 ? K - 1 : lo);
  const double t0 = ts[i - 1], t1 = ts[i];
  const double os = o.output_scale[inst * K + i];
  const size_t r0 = inst * K + (i - 1), r1 = inst * K + i;
  const bool filter = c.a->cfg.strategy == PDQ_FILTER;
  const double* m0p = (filter ? o.mean : o.filt_mean) + r0 * ms;
  const double* L0p = (filter ? o.chol : o.filt_chol) + r0 * cs;
  double *X0m = c.slot(0) + L.s_m, *X0L = c.slot(0) + L.s_L;
  double *Xm = c.slot(1) + L.s_m, *XL = c.slot(1) + L.s_L;
  double *Zm = c.slot(2) + L.s_m, *ZL = c.slot(2) + L.s_L;
  double *P = c.W(L.w_P), *PI = c.W(L.w_PI), *big = c.W(L.w_big), *tmp = c.W(L.w_T1);
  double *Ym = c.W(L.w_Pm), *YL = c.W(L.w_PL);
  double *bA = c.W(L.w_bA), *bm = c.W(L.w_bm), *bL = c.W(L.w_bL), *btl = c.W(L.w_btl), *bto = c.W(L.w_bto);
  g_copy(c.g, X0m, m0p, ms);
  g_copy(c.g, X0L, L0p, cs);
  c.g.sync();
  precon(c, t - t0, P, PI);
  c.g.sync();
  const double s0 = noise_scale(c, t - t0, os);
  if (filter) {
    marginalise_prior(c, P, PI, s0, X0m, X0L, Ym, YL, big);
  } else {
    revert_prior(c, P, PI, s0, X0m, X0L, Xm, XL, bA, bm, bL, btl, bto, big, tmp);
    c.g.sync();
    precon(c, t1 - t, P, PI);
    c.g.sync();
    const double s1 = noise_scale(c, t1 - t, os);
    revert_prior(c, P, PI, s1, Xm, XL, Zm, ZL, bA, bm, bL, btl, bto, big, tmp);
    g_copy(c.g, X0m, o.mean + r1 * ms, ms);
    g_copy(c.g, X0L, o.chol + r1 * cs, cs);
    c.g.sync();
    marginalise_cond(c, bA, bm, bL, btl, bto, X0m, X0L, Ym, YL, big);
  }
  c.g.sync();
  g_copy(c.g, o.mean_out + q * ms, Ym, ms);
  g_copy(c.g, o.chol_out + q * cs, YL, cs);
  c.g.sync();
}

#ifndef PDQ_HOST_EMU
// Workgroup set-up shared by the cooperative kernels: the IWP matrices (and
// the exponential prior's drift rows) in LDS, and the context of this lane's
// group.
template <int G>
__device__ Ctx<G> coop_setup(const KArgs& a, double* lds) {
  const Layout& L = a.lay;
  const int nc = L.nc;
  double* sA = lds;
  double* sQ = lds + nc * nc;
  double* sD = lds + 2 * nc * nc;
  for (int i = threadIdx.x; i < nc * nc; i += blockDim.x) {
    sA[i] = a.A[i];
    sQ[i] = a.Q[i];
  }
  if (L.expo)
    for (int i = threadIdx.x; i < L.d * nc; i += blockDim.x) sD[i] = a.cfg.prior_drift[i];
  __syncthreads();
  const int grp = threadIdx.x / G;
  Ctx<G> c;
  c.g.l = threadIdx.x % G;
  c.a = &a;
  c.L = &L;
  c.nc = L.nc;
  c.mc = L.mc;
  c.mcp = L.mcp;
  c.ko = L.ko;
  c.N = L.N;
  c.d = L.d;
  c.dense = a.cfg.factorisation == PDQ_DENSE;
  c.ws = lds + L.wg_shared + grp * L.inst_size;
  c.prm_ = c.ws + L.w_prm;
  c.expo = L.expo;
  c.sD_ = sD;
  // the exponential prior's transition lives in the instance's workspace
  c.sA_ = L.expo ? c.ws + L.w_xA : sA;
  c.sQ_ = L.expo ? c.ws + L.w_xL : sQ;
  return c;
}

template <int G>
__global__ void __launch_bounds__(64) offgrid_coop_kernel(OffgridCoopArgs o) {
  double* const lds = pdq_coop_lds;
  const Ctx<G> c = coop_setup<G>(o.k, lds);
  const size_t total = (size_t)o.k.n_inst * o.n_query;
  const size_t groups = (size_t)gridDim.x * (64 / G);
  for (size_t q = (size_t)blockIdx.x * (64 / G) + threadIdx.x / G; q < total; q += groups) offgrid_query(c, o, q);
}

template <int G>
__global__ void __launch_bounds__(64) solve_kernel(KArgs args) {
  double* const lds = pdq_coop_lds;
  const KArgs& a = args;
  const Ctx<G> c = coop_setup<G>(a, lds);
  for (;;) {
    int inst = 0;
    if (c.g.l == 0) inst = atomicAdd(a.queue, 1);
    inst = c.g.bcast(inst);
    if (inst >= a.n_inst) break;
    if (a.mode == MODE_LML) lml_instance(c, inst);
    else solve_instance(c, inst);
  }
}


// Initial Taylor coefficients with forward-mode tangents d/d params and
// d/d y0 (jet_expansion_algorithms.py:49-103 for the registry); one thread
// per instance.  out: n_inst x (N d) x (1 + ntan).
__global__ void taylor_tangent_kernel(int vf, int d, int np, int n_inst, int num, int ntan, const double* params,
                                      const double* y0, const double* t0p, double* out) {
  const int inst = blockIdx.x * blockDim.x + threadIdx.x;
  if (inst >= n_inst) return;
  const double t0 = t0p[0];
  Dual prmD[kMaxParams];
  Dual cc[64];
  for (int q = 0; q < np; ++q) {
    prmD[q] = dconst(params[(size_t)inst * np + q]);
    if (q < ntan) prmD[q].g[q] = 1.0;
  }
  for (int k = 0; k < d; ++k) {
    cc[k] = dconst(y0[(size_t)inst * d + k]);
    if (np + k < ntan) cc[k].g[np + k] = 1.0;
  }
  for (int k = 0; k < num; ++k)
    for (int j = 0; j < d; ++j) cc[(k + 1) * d + j] = taylor_rhs(vf, d, cc, prmD, t0, k, j) / (double)(k + 1);
  const int N = num + 1;
  for (int i = 0; i < N; ++i) {
    const double f = factorial_d(i);
    for (int k = 0; k < d; ++k) {
      const Dual v = f * cc[i * d + k];
      for (int q = 0; q <= ntan; ++q)
        out[(((size_t)inst * N + i) * d + k) * (1 + ntan) + q] = q == 0 ? v.v : v.g[q - 1];
    }
  }
}

__global__ void zero_queue(int* p) {
  if (threadIdx.x == 0) p[0] = 0;
}

int launch_taylor_tangents(int vf, int d, int np, int n_inst, int num, int ntan, const double* params,
                           const double* y0, const double* t0, double* out, hipStream_t st) {
  hipLaunchKernelGGL(taylor_tangent_kernel, dim3((n_inst + 63) / 64), dim3(64), 0, st, vf, d, np, n_inst, num, ntan,
                     params, y0, t0, out);
  PDQ_HIP_TRY(hipGetLastError());
  return PDQ_OK;
}

size_t coop_lds_bytes(const Layout& L, int G) { return sizeof(double) * (size_t)(L.wg_shared + (64 / G) * L.inst_size); }

template <int G>
int launch_coop_g(KArgs& a, hipStream_t st) {
  const size_t lds = coop_lds_bytes(a.lay, G);
  auto kern = solve_kernel<G>;
  if (hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) {
    (void)hipGetLastError();
    return host_error(PDQ_ERR_UNSUPPORTED,
                      "lanes_per_instance = %d puts %d instances and %zu bytes of LDS in a workgroup, more than the "
                      "device allows; take more lanes per instance",
                      G, 64 / G, lds);
  }
  int per_cu = 0;
  PDQ_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64, lds));
  if (per_cu < 1) return host_error(PDQ_ERR_UNSUPPORTED, "instance does not fit into LDS (%zu bytes)", lds);
  const int ipw = 64 / G;
  const long need = ((long)a.n_inst + ipw - 1) / ipw;
  const long cap = (long)per_cu * num_cus();
  const int grid = (int)std::max<long>(1, std::min(need, cap));
  hipLaunchKernelGGL(zero_queue, dim3(1), dim3(64), 0, st, a.queue);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(64), lds, st, a);
  PDQ_HIP_TRY(hipGetLastError());
  return PDQ_OK;
}

template <int G>
int launch_offgrid_coop_g(OffgridCoopArgs& o, hipStream_t st) {
  const size_t lds = coop_lds_bytes(o.k.lay, G);
  auto kern = offgrid_coop_kernel<G>;
  if (hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) {
    (void)hipGetLastError();
    return host_error(PDQ_ERR_UNSUPPORTED, "%d queries of %zu bytes of LDS do not fit a workgroup", 64 / G, lds);
  }
  int per_cu = 0;
  PDQ_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64, lds));
  if (per_cu < 1) return host_error(PDQ_ERR_UNSUPPORTED, "a query does not fit into LDS (%zu bytes)", lds);
  const long need = ((long)o.k.n_inst * o.n_query + 64 / G - 1) / (64 / G);
  const long cap = (long)per_cu * num_cus() * 4;
  const int grid = (int)std::max<long>(1, std::min(need, cap));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(64), lds, st, o);
  PDQ_HIP_TRY(hipGetLastError());
  return PDQ_OK;
}

int launch_offgrid_coop(OffgridCoopArgs& o, int G, hipStream_t st) {
  switch (G) {
    case 4: return launch_offgrid_coop_g<4>(o, st);
    case 8: return launch_offgrid_coop_g<8>(o, st);
    case 16: return launch_offgrid_coop_g<16>(o, st);
    case 32: return launch_offgrid_coop_g<32>(o, st);
    case 64: return launch_offgrid_coop_g<64>(o, st);
  }
  return host_error(PDQ_ERR_INVALID, "lanes per query must be one of 4, 8, 16, 32, 64 (got %d)", G);
}

int launch_coop(KArgs& a, int G, hipStream_t st) {
  switch (G) {
    case 4: return launch_coop_g<4>(a, st);
    case 8: return launch_coop_g<8>(a, st);
    case 16: return launch_coop_g<16>(a, st);
    case 32: return launch_coop_g<32>(a, st);
    case 64: return launch_coop_g<64>(a, st);
  }
  return host_error(PDQ_ERR_INVALID, "lanes_per_instance must be one of 4, 8, 16, 32, 64 (got %d)", G);
}


#else
// CPU build (G = 1); ws must hold L.inst_size doubles and shared must hold
// L.wg_shared doubles.
inline Ctx<1> emu_ctx(const KArgs& a, double* shared, double* ws) {
  const Layout& L = a.lay;
  const int nc = L.nc;
  for (int i = 0; i < nc * nc; ++i) {
    shared[i] = a.A[i];
    shared[nc * nc + i] = a.Q[i];
  }
  if (L.expo)
    for (int i = 0; i < L.d * nc; ++i) shared[2 * nc * nc + i] = a.cfg.prior_drift[i];
  Ctx<1> c;
  c.g.l = 0;
  c.a = &a;
  c.L = &L;
  c.nc = L.nc;
  c.mc = L.mc;
  c.mcp = L.mcp;
  c.ko = L.ko;
  c.N = L.N;
  c.d = L.d;
  c.dense = a.cfg.factorisation == PDQ_DENSE;
  c.ws = ws;
  c.prm_ = ws + L.w_prm;
  c.expo = L.expo;
  c.sD_ = shared + 2 * nc * nc;
  c.sA_ = L.expo ? ws + L.w_xA : shared;
  c.sQ_ = L.expo ? ws + L.w_xL : shared + nc * nc;
  return c;
}

inline void emu_solve_instance(const KArgs& a, int inst, double* shared, double* ws) {
  const Ctx<1> c = emu_ctx(a, shared, ws);
  if (a.mode == MODE_LML) lml_instance(c, inst);
  else solve_instance(c, inst);
}

inline void emu_offgrid_query(const OffgridCoopArgs& o, size_t q, double* shared, double* ws) {
  const Ctx<1> c = emu_ctx(o.k, shared, ws);
  offgrid_query(c, o, q);
}
#endif
}  // namespace pdq
