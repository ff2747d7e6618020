// Host-side dense linear algebra for the operator set-up: LU with partial
// pivoting and the eigen-decomposition of a real matrix whose eigenvalues are
// real (Householder Hessenberg reduction, shifted QR with Givens rotations,
// eigenvectors of the triangular Schur factor by back-substitution).  Sizes are
// the wall-normal half systems (n <= 129); matrices are row-major vectors.
#pragma once
#include <algorithm>
#include <cmath>
#include <vector>

namespace chf {
namespace la {

using Mat = std::vector<double>;  // row-major n x n

// In-place LU with partial pivoting; returns false if singular.
inline bool lu(Mat& a, std::vector<int>& piv, int n) {
  piv.resize(n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::fabs(a[(size_t)k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(a[(size_t)i * n + k]) > best) { best = std::fabs(a[(size_t)i * n + k]); p = i; }
    if (best == 0.0) return false;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(a[(size_t)k * n + j], a[(size_t)p * n + j]);
    const double inv = 1.0 / a[(size_t)k * n + k];
    for (int i = k + 1; i < n; ++i) {
      const double f = a[(size_t)i * n + k] * inv;
      a[(size_t)i * n + k] = f;
      if (f != 0.0)
        for (int j = k + 1; j < n; ++j) a[(size_t)i * n + j] -= f * a[(size_t)k * n + j];
    }
  }
  return true;
}

// Solves A X = B for m right-hand sides (B row-major n x m, overwritten).
inline void lu_solve(const Mat& lu_, const std::vector<int>& piv, int n, Mat& b, int m) {
  for (int k = 0; k < n; ++k)
    if (piv[k] != k)
      for (int j = 0; j < m; ++j) std::swap(b[(size_t)k * m + j], b[(size_t)piv[k] * m + j]);
  for (int i = 1; i < n; ++i)
    for (int k = 0; k < i; ++k) {
      const double f = lu_[(size_t)i * n + k];
      if (f != 0.0)
        for (int j = 0; j < m; ++j) b[(size_t)i * m + j] -= f * b[(size_t)k * m + j];
    }
  for (int i = n - 1; i >= 0; --i) {
    const double d = 1.0 / lu_[(size_t)i * n + i];
    for (int j = 0; j < m; ++j) {
      double s = b[(size_t)i * m + j];
      for (int k = i + 1; k < n; ++k) s -= lu_[(size_t)i * n + k] * b[(size_t)k * m + j];
      b[(size_t)i * m + j] = s * d;
    }
  }
}

// Inverse of A (n x n); false if singular.
inline bool inverse(const Mat& a, int n, Mat& out) {
  Mat f = a;
  std::vector<int> piv;
  if (!lu(f, piv, n)) return false;
  out.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) out[(size_t)i * n + i] = 1.0;
  lu_solve(f, piv, n, out, n);
  return true;
}

inline Mat matmul(const Mat& a, const Mat& b, int n, int k, int m) {  // (n x k)(k x m)
  Mat c((size_t)n * m, 0.0);
  for (int i = 0; i < n; ++i)
    for (int p = 0; p < k; ++p) {
      const double f = a[(size_t)i * k + p];
      if (f != 0.0)
        for (int j = 0; j < m; ++j) c[(size_t)i * m + j] += f * b[(size_t)p * m + j];
    }
  return c;
}

// Eigen-decomposition A = V diag(lam) V^-1 of a real n x n matrix with real,
// distinct eigenvalues.  Returns false if the QR iteration does not reduce A
// to triangular form (complex pairs, no convergence) or an eigenvector
// degenerates; V's columns are normalised to unit 2-norm.
inline bool eig_real(const Mat& A, int n, std::vector<double>& lam, Mat& V) {
  Mat H = A;
  Mat Q((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) Q[(size_t)i * n + i] = 1.0;
  auto h = [&](int i, int j) -> double& { return H[(size_t)i * n + j]; };
  auto q = [&](int i, int j) -> double& { return Q[(size_t)i * n + j]; };
  // Householder reduction to upper Hessenberg form, H = Q^T A Q
  std::vector<double> v(n);
  for (int k = 0; k < n - 2; ++k) {
    double alpha = 0.0;
    for (int i = k + 1; i < n; ++i) alpha += h(i, k) * h(i, k);
    alpha = std::sqrt(alpha);
    if (alpha == 0.0) continue;
    if (h(k + 1, k) > 0) alpha = -alpha;
    double vn = 0.0;
    for (int i = 0; i < n; ++i) v[i] = 0.0;
    v[k + 1] = h(k + 1, k) - alpha;
    for (int i = k + 2; i < n; ++i) v[i] = h(i, k);
    for (int i = k + 1; i < n; ++i) vn += v[i] * v[i];
    if (vn == 0.0) continue;
    const double beta = 2.0 / vn;
    for (int j = 0; j < n; ++j) {  // H = (I - beta v v^T) H
      double s = 0.0;
      for (int i = k + 1; i < n; ++i) s += v[i] * h(i, j);
      s *= beta;
      for (int i = k + 1; i < n; ++i) h(i, j) -= s * v[i];
    }
    for (int i = 0; i < n; ++i) {  // H = H (I - beta v v^T), Q = Q (I - beta v v^T)
      double s = 0.0, t = 0.0;
      for (int j = k + 1; j < n; ++j) {
        s += h(i, j) * v[j];
        t += q(i, j) * v[j];
      }
      s *= beta;
      t *= beta;
      for (int j = k + 1; j < n; ++j) {
        h(i, j) -= s * v[j];
        q(i, j) -= t * v[j];
      }
    }
    for (int i = k + 2; i < n; ++i) h(i, k) = 0.0;
  }
  // Shifted QR on the active block [lo, hi] with Givens rotations; deflate at
  // the bottom.  Shift: the eigenvalue of the trailing 2 x 2 block closest to
  // its last diagonal entry (its real part if the pair is complex), with an
  // exceptional shift every 11 iterations without deflation.
  const double eps = 2.220446049250313e-16;
  int hi = n - 1, iter = 0, total = 0;
  std::vector<double> cs(n), sn(n);
  while (hi > 0) {
    int lo = hi;
    while (lo > 0) {
      const double s = std::fabs(h(lo - 1, lo - 1)) + std::fabs(h(lo, lo));
      if (std::fabs(h(lo, lo - 1)) <= eps * (s > 0 ? s : 1.0)) {
        h(lo, lo - 1) = 0.0;
        break;
      }
      --lo;
    }
    if (lo == hi) {
      --hi;
      iter = 0;
      continue;
    }
    if (++total > 100 * n) return false;
    ++iter;
    const double a = h(hi - 1, hi - 1), b = h(hi - 1, hi), c = h(hi, hi - 1), d = h(hi, hi);
    const double tr = 0.5 * (a - d), disc = tr * tr + b * c;
    double mu;
    if (disc >= 0) {
      const double r = std::sqrt(disc);
      mu = d - (b * c) / (tr + (tr >= 0 ? r : -r) + (tr == 0 && r == 0 ? 1e-300 : 0.0));
    } else {
      mu = 0.5 * (a + d);
    }
    if (iter % 11 == 0) mu = d + std::fabs(h(hi, hi - 1));
    for (int i = lo; i <= hi; ++i) h(i, i) -= mu;
    for (int k = lo; k < hi; ++k) {  // QR: rotate rows k, k+1
      const double x = h(k, k), y = h(k + 1, k);
      const double r = std::hypot(x, y);
      const double cc = r == 0 ? 1.0 : x / r, ss = r == 0 ? 0.0 : y / r;
      cs[k] = cc;
      sn[k] = ss;
      for (int j = k; j < n; ++j) {
        const double t1 = h(k, j), t2 = h(k + 1, j);
        h(k, j) = cc * t1 + ss * t2;
        h(k + 1, j) = -ss * t1 + cc * t2;
      }
    }
    for (int k = lo; k < hi; ++k) {  // RQ: rotate columns k, k+1 (and Q)
      const double cc = cs[k], ss = sn[k];
      for (int i = 0; i <= std::min(k + 2, hi); ++i) {
        const double t1 = h(i, k), t2 = h(i, k + 1);
        h(i, k) = cc * t1 + ss * t2;
        h(i, k + 1) = -ss * t1 + cc * t2;
      }
      for (int i = 0; i < n; ++i) {
        const double t1 = q(i, k), t2 = q(i, k + 1);
        q(i, k) = cc * t1 + ss * t2;
        q(i, k + 1) = -ss * t1 + cc * t2;
      }
    }
    for (int i = lo; i <= hi; ++i) h(i, i) += mu;
  }
  // rows above the active window were not rotated by the column updates of
  // later deflated blocks; the row updates above cover j >= k, so H is the
  // upper-triangular Schur factor T with A = Q T Q^T.
  lam.resize(n);
  for (int i = 0; i < n; ++i) lam[i] = h(i, i);
  // eigenvectors of T: T y = lam_i y, y_i = 1, y_j = 0 (j > i)
  Mat Y((size_t)n * n, 0.0);
  double tnorm = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) tnorm = std::max(tnorm, std::fabs(h(i, j)));
  for (int i = 0; i < n; ++i) {
    Y[(size_t)i * n + i] = 1.0;
    for (int j = i - 1; j >= 0; --j) {
      double s = 0.0;
      for (int k = j + 1; k <= i; ++k) s += h(j, k) * Y[(size_t)k * n + i];
      double den = h(j, j) - lam[i];
      if (std::fabs(den) < eps * tnorm) den = (den >= 0 ? 1.0 : -1.0) * eps * tnorm;
      Y[(size_t)j * n + i] = -s / den;
    }
  }
  V = matmul(Q, Y, n, n, n);
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int r = 0; r < n; ++r) s += V[(size_t)r * n + i] * V[(size_t)r * n + i];
    s = std::sqrt(s);
    if (!(s > 0) || !std::isfinite(s)) return false;
    for (int r = 0; r < n; ++r) V[(size_t)r * n + i] /= s;
  }
  return true;
}

}  // namespace la
}  // namespace chf
