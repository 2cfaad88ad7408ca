// Small dense linear algebra on the host: Hermitian eigen-decomposition of the
// projected (Krylov / Loewdin) matrices, which are at most a few hundred wide.
//
// Method: Householder reduction of the complex Hermitian matrix to a complex
// Hermitian tridiagonal one, a diagonal unitary that makes the off-diagonal
// real, then the implicit-shift QL iteration on the real symmetric tridiagonal
// matrix with accumulated rotations (Wilkinson & Reinsch, Handbook for Automatic
// Computation vol. II, procedures tred/tql2).
#include <algorithm>
#include <cmath>
#include <complex>
#include <numeric>
#include <vector>

#include "qs_internal.h"

using cd = std::complex<double>;

namespace {

// A: column-major n*n (A[i + j*n] = row i, col j), Hermitian; only used as input.
// Returns w ascending and Z (column-major, column j = eigenvector j).
bool heev_impl(int n, const cd *Ain, double *w, cd *Z) {
  if (n <= 0) return true;
  std::vector<cd> A(Ain, Ain + (size_t)n * n);
  auto at = [&](int i, int j) -> cd & { return A[i + (size_t)j * n]; };
  std::vector<cd> Q((size_t)n * n, cd(0));
  for (int i = 0; i < n; ++i) Q[i + (size_t)i * n] = 1.0;
  auto q = [&](int i, int j) -> cd & { return Q[i + (size_t)j * n]; };
  std::vector<cd> v(n), p(n), wv(n);
  // Householder: zero column k below the subdiagonal
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;
    double alpha = 0.0;
    for (int i = 0; i < m; ++i) alpha += std::norm(at(k + 1 + i, k));
    alpha = std::sqrt(alpha);
    if (alpha == 0.0) continue;
    cd x0 = at(k + 1, k);
    cd phase = std::abs(x0) > 0 ? x0 / std::abs(x0) : cd(1.0);
    for (int i = 0; i < m; ++i) v[i] = at(k + 1 + i, k);
    v[0] += phase * alpha;
    double vnorm2 = 0.0;
    for (int i = 0; i < m; ++i) vnorm2 += std::norm(v[i]);
    if (vnorm2 == 0.0) continue;
    const double beta = 2.0 / vnorm2;
    // p = beta * A22 v
    for (int i = 0; i < m; ++i) {
      cd s = 0;
      for (int j = 0; j < m; ++j) s += at(k + 1 + i, k + 1 + j) * v[j];
      p[i] = beta * s;
    }
    cd K = 0;
    for (int i = 0; i < m; ++i) K += std::conj(v[i]) * p[i];
    K *= 0.5 * beta;
    for (int i = 0; i < m; ++i) wv[i] = p[i] - K * v[i];
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i)
        at(k + 1 + i, k + 1 + j) -= v[i] * std::conj(wv[j]) + wv[i] * std::conj(v[j]);
    const cd sub = -phase * alpha;
    at(k + 1, k) = sub;
    at(k, k + 1) = std::conj(sub);
    for (int i = 1; i < m; ++i) {
      at(k + 1 + i, k) = 0;
      at(k, k + 1 + i) = 0;
    }
    // Q <- Q H on columns k+1..n-1
    for (int r = 0; r < n; ++r) {
      cd s = 0;
      for (int j = 0; j < m; ++j) s += q(r, k + 1 + j) * v[j];
      s *= beta;
      for (int j = 0; j < m; ++j) q(r, k + 1 + j) -= s * std::conj(v[j]);
    }
  }
  // real tridiagonal via diagonal phases
  std::vector<double> d(n), e(n, 0.0);
  std::vector<cd> D(n, cd(1.0));
  for (int i = 0; i < n; ++i) d[i] = at(i, i).real();
  for (int i = 0; i + 1 < n; ++i) {
    const cd s = at(i + 1, i);
    const double a = std::abs(s);
    e[i] = a;
    D[i + 1] = a > 0 ? D[i] * (s / a) : D[i];
  }
  // implicit QL with shifts; Zr accumulates rotations (row-major n*n)
  std::vector<double> Zr((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) Zr[(size_t)i * n + i] = 1.0;
  const double eps = std::numeric_limits<double>::epsilon();
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    while (true) {
      int m = l;
      for (; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= eps * dd) break;
      }
      if (m == l) break;
      if (++iter > 200) return false;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
      double s = 1.0, c = 1.0, pp = 0.0;
      bool early = false;
      for (int i = m - 1; i >= l; --i) {
        double f = s * e[i];
        const double b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= pp;
          e[m] = 0.0;
          early = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - pp;
        r = (d[i] - g) * s + 2.0 * c * b;
        pp = s * r;
        d[i + 1] = g + pp;
        g = c * r - b;
        for (int k2 = 0; k2 < n; ++k2) {
          f = Zr[(size_t)k2 * n + i + 1];
          Zr[(size_t)k2 * n + i + 1] = s * Zr[(size_t)k2 * n + i] + c * f;
          Zr[(size_t)k2 * n + i] = c * Zr[(size_t)k2 * n + i] - s * f;
        }
      }
      if (early) continue;
      d[l] -= pp;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  // sort ascending, eigenvectors = (Q D) Zr
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return d[a] < d[b]; });
  for (int jj = 0; jj < n; ++jj) {
    const int j = ord[jj];
    w[jj] = d[j];
    for (int r = 0; r < n; ++r) {
      cd s = 0;
      for (int k2 = 0; k2 < n; ++k2) s += q(r, k2) * D[k2] * Zr[(size_t)k2 * n + j];
      Z[r + (size_t)jj * n] = s;
    }
  }
  return true;
}

}  // namespace

void qs_heev(int n, const double *A, double *w, double *Z) {
  heev_impl(n, reinterpret_cast<const cd *>(A), w, reinterpret_cast<cd *>(Z));
}

// ---------------------------------------------------------------------------
// real symmetric tridiagonal T (diagonal d[0..m-1], off-diagonal e[0..m-2])
// ---------------------------------------------------------------------------
namespace {

// number of eigenvalues of T below x (Sturm sequence of the LDL^T pivots)
int sturm_count(int m, const double *d, const double *e, double x) {
  int cnt = 0;
  double q = d[0] - x;
  if (q < 0) ++cnt;
  for (int i = 1; i < m; ++i) {
    const double qq = q != 0.0 ? q : std::numeric_limits<double>::epsilon() * (std::fabs(e[i - 1]) + 1e-300);
    q = d[i] - x - e[i - 1] * e[i - 1] / qq;
    if (q < 0) ++cnt;
  }
  return cnt;
}

// k-th smallest eigenvalue (k = 0..m-1) by bisection on the Gershgorin interval
double kth_eigenvalue(int m, const double *d, const double *e, int k) {
  double lo = d[0], hi = d[0];
  for (int i = 0; i < m; ++i) {
    const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
  }
  const double span = std::max(hi - lo, 1e-300);
  lo -= 1e-12 * span;
  hi += 1e-12 * span;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (sturm_count(m, d, e, mid) > k) hi = mid; else lo = mid;
  }
  return 0.5 * (lo + hi);
}

}  // namespace

// Lowest eigenpair of T: eigenvalue by Sturm bisection, eigenvector by inverse
// iteration with the tridiagonal LU (Thomas algorithm, partial pivoting).
extern "C" qs_status qs_tridiag_lowest_host(int m, const double *d, const double *e, double *E, double *z) {
  if (m < 1 || !d || !E || !z || (m > 1 && !e)) return QS_EINVAL;
  const double lam = kth_eigenvalue(m, d, e, 0);
  *E = lam;
  if (m == 1) {
    z[0] = 1.0;
    return QS_OK;
  }
  double scale = 0.0;
  for (int i = 0; i < m; ++i) scale = std::max(scale, std::fabs(d[i]) + (i + 1 < m ? std::fabs(e[i]) : 0.0));
  const double shift = lam - 1e-14 * std::max(scale, 1e-300);
  // LU of (T - shift) with partial pivoting: rows i, i+1 may swap; U has 3 diagonals
  std::vector<double> u0(m), u1(m, 0.0), u2(m, 0.0), l(m, 0.0);
  std::vector<int> swp(m, 0);
  double a = d[0] - shift, bnext = e[0];  // current row: diagonal a, super-diagonal bnext
  double c = 0.0;                            // second super-diagonal carried into the current row
  for (int i = 0; i < m - 1; ++i) {
    const double sub = e[i];                  // T[i+1][i]
    const double dn = d[i + 1] - shift;       // T[i+1][i+1]
    const double en = i + 2 < m ? e[i + 1] : 0.0;
    if (std::fabs(sub) > std::fabs(a)) {
      // swap rows i and i+1
      swp[i] = 1;
      u0[i] = sub;
      u1[i] = dn;
      u2[i] = en;
      const double f = a / sub;
      l[i] = f;
      a = bnext - f * dn;
      bnext = c - f * en;
      c = 0.0;
    } else {
      u0[i] = a != 0.0 ? a : 1e-300;
      u1[i] = bnext;
      u2[i] = c;
      const double f = sub / u0[i];
      l[i] = f;
      a = dn - f * bnext;
      bnext = en - f * c;
      c = 0.0;
    }
  }
  u0[m - 1] = a != 0.0 ? a : 1e-300;
  std::vector<double> x(m, 1.0), y(m);
  for (int iter = 0; iter < 4; ++iter) {
    // forward: apply the row swaps and L
    for (int i = 0; i < m; ++i) y[i] = x[i];
    for (int i = 0; i < m - 1; ++i) {
      if (swp[i]) std::swap(y[i], y[i + 1]);
      y[i + 1] -= l[i] * y[i];
    }
    // back substitution with U
    for (int i = m - 1; i >= 0; --i) {
      double s = y[i];
      if (i + 1 < m) s -= u1[i] * x[i + 1];
      if (i + 2 < m) s -= u2[i] * x[i + 2];
      x[i] = s / u0[i];
    }
    double nrm = 0.0;
    for (int i = 0; i < m; ++i) nrm += x[i] * x[i];
    nrm = std::sqrt(nrm);
    for (int i = 0; i < m; ++i) x[i] /= nrm;
  }
  // fix the sign: largest component positive
  int im = 0;
  for (int i = 1; i < m; ++i)
    if (std::fabs(x[i]) > std::fabs(x[im])) im = i;
  const double sg = x[im] < 0 ? -1.0 : 1.0;
  for (int i = 0; i < m; ++i) z[i] = sg * x[i];
  return QS_OK;
}

// Smallest and largest eigenvalue of T (pole range of a Lanczos table).
extern "C" qs_status qs_tridiag_extremes_host(int m, const double *d, const double *e, double *lo, double *hi) {
  if (m < 1 || !d || !lo || !hi || (m > 1 && !e)) return QS_EINVAL;
  *lo = kth_eigenvalue(m, d, e, 0);
  *hi = kth_eigenvalue(m, d, e, m - 1);
  return QS_OK;
}

extern "C" qs_status qs_heev_host(int n, const double *A, double *w, double *Z) {
  if (n < 0 || (n > 0 && (!A || !w || !Z))) return QS_EINVAL;
  if (!heev_impl(n, reinterpret_cast<const cd *>(A), w, reinterpret_cast<cd *>(Z))) return QS_ENOTCONVERGED;
  return QS_OK;
}
