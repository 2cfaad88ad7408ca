// Small dense linear algebra on the host: Hermitian eigen-decomposition of the
// projected (Krylov / Loewdin) matrices, which are at most a few hundred wide.
//
// Method: Householder reduction of the complex Hermitian matrix to a complex
// Hermitian tridiagonal one (lower triangle only, column access), a diagonal
// unitary that makes the off-diagonal real, then the implicit-shift QL
// iteration on the real symmetric tridiagonal matrix (Wilkinson & Reinsch,
// Handbook for Automatic Computation vol. II, procedures tred/tql2). When only
// a few eigenvectors are wanted, the QL rotations are recorded and replayed
// on those unit vectors, and only they are transformed back; otherwise the
// rotations are accumulated into a column-major matrix.
#include <algorithm>
#include <cmath>
#include <complex>
#include <numeric>
#include <vector>

#include "qs_internal.h"

using cd = std::complex<double>;

namespace {

// sqrt(a^2 + b^2) without std::hypot's cost (it dominated the QL sweeps of
// the dense spectra: one call per plane rotation, ~1.5 n^2 rotations). The
// plain formula is exact to a few ulps while a^2 + b^2 neither overflows nor
// underflows; outside that range std::hypot takes over.
inline double hypot_fast(double a, double b) {
  const double m = std::fmax(std::fabs(a), std::fabs(b));
  if (m > 1e150 || m < 1e-150) return m == 0.0 ? 0.0 : std::hypot(a, b);
  return std::sqrt(a * a + b * b);
}

// A: column-major n*n (A[i + j*n] = row i, col j), Hermitian; only its lower
// triangle is read. Returns all eigenvalues ascending in w and the eigenvectors
// of the nvec lowest in Z (column-major n x nvec).
// Implicit QL with Wilkinson-type shifts on the symmetric tridiagonal matrix
// (d[0..n-1], e[0..n-2]; e has n entries, e[n-1] ignored): d receives the
// eigenvalues (unsorted), e is destroyed. Every plane rotation acting on
// components (i, i+1) is handed to rot(i, c, s) in the order it is made: with
// Z = G_1 G_2 ... G_K the eigenvector matrix, the row update
// (r_i, r_{i+1}) <- (c r_i - s r_{i+1}, s r_i + c r_{i+1}) turns any block R
// into Z^T R. False when an eigenvalue takes more than 200 iterations.
template <class F>
bool ql_rotations(int n, double *d, double *e, F &&rot) {
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
      double r = hypot_fast(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
      double s = 1.0, c = 1.0, pp = 0.0;
      bool early = false;
      for (int i = m - 1; i >= l; --i) {
        double f = s * e[i];
        const double b = c * e[i];
        r = hypot_fast(f, g);
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
        rot(i, c, s);
      }
      if (early) continue;
      d[l] -= pp;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  return true;
}

bool heev_impl(int n, const cd *Ain, double *w, cd *Z, int nvec) {
  if (n <= 0) return true;
  nvec = std::max(0, std::min(nvec, n));
  const size_t nn = (size_t)n * n;
  std::vector<double> ar(nn), ai(nn);
  for (size_t t = 0; t < nn; ++t) {
    ar[t] = Ain[t].real();
    ai[t] = Ain[t].imag();
  }
  std::vector<double> bet(n, 0.0), er(n, 0.0), ei(n, 0.0);
  std::vector<double> pr(n), pi(n), wr(n), wi(n);
  // Householder step k: v (stored over column k, rows k+1..n-1) maps the column
  // below the diagonal onto -phase*alpha e1; A22 <- H A22 H = A22 - v w^+ - w v^+
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;
    double *xr = &ar[(size_t)(k + 1) + (size_t)k * n], *xi = &ai[(size_t)(k + 1) + (size_t)k * n];
    double a2 = 0.0;
    for (int i = 0; i < m; ++i) a2 += xr[i] * xr[i] + xi[i] * xi[i];
    if (a2 == 0.0) continue;  // already reduced: e_k = 0, H_k = I
    const double alpha = std::sqrt(a2);
    const double ax0 = std::hypot(xr[0], xi[0]);
    const double phr = ax0 > 0 ? xr[0] / ax0 : 1.0, phi = ax0 > 0 ? xi[0] / ax0 : 0.0;
    er[k] = -phr * alpha;
    ei[k] = -phi * alpha;
    xr[0] += phr * alpha;
    xi[0] += phi * alpha;
    const double b = 1.0 / (a2 + alpha * ax0);  // 2 / |v|^2
    bet[k] = b;
    // p = b A22 v from the lower triangle: column j adds A_ij v_j to p_i (i > j)
    // and conj(A_ij) v_i to p_j
    std::fill(pr.begin(), pr.begin() + m, 0.0);
    std::fill(pi.begin(), pi.begin() + m, 0.0);
    for (int jj = 0; jj < m; ++jj) {
      const size_t cj = (size_t)(k + 1 + jj) * n + (k + 1);
      const double *cr = &ar[cj], *ci = &ai[cj];
      const double vr = xr[jj], vi = xi[jj];
      double sr = cr[jj] * vr, si = cr[jj] * vi;
      for (int ii = jj + 1; ii < m; ++ii) {
        const double r = cr[ii], im = ci[ii];
        pr[ii] += r * vr - im * vi;
        pi[ii] += r * vi + im * vr;
        sr += r * xr[ii] + im * xi[ii];
        si += r * xi[ii] - im * xr[ii];
      }
      pr[jj] += sr;
      pi[jj] += si;
    }
    double kr = 0.0, ki = 0.0;  // v^+ p (real up to rounding)
    for (int i = 0; i < m; ++i) {
      pr[i] *= b;
      pi[i] *= b;
      kr += xr[i] * pr[i] + xi[i] * pi[i];
      ki += xr[i] * pi[i] - xi[i] * pr[i];
    }
    kr *= 0.5 * b;
    ki *= 0.5 * b;
    for (int i = 0; i < m; ++i) {
      wr[i] = pr[i] - (kr * xr[i] - ki * xi[i]);
      wi[i] = pi[i] - (kr * xi[i] + ki * xr[i]);
    }
    for (int jj = 0; jj < m; ++jj) {
      const size_t cj = (size_t)(k + 1 + jj) * n + (k + 1);
      double *cr = &ar[cj], *ci = &ai[cj];
      const double vr = xr[jj], vi = xi[jj], w_r = wr[jj], w_i = wi[jj];
      for (int ii = jj; ii < m; ++ii) {
        cr[ii] -= (xr[ii] * w_r + xi[ii] * w_i) + (wr[ii] * vr + wi[ii] * vi);
        ci[ii] -= (xi[ii] * w_r - xr[ii] * w_i) + (wi[ii] * vr - wr[ii] * vi);
      }
      ci[jj] = 0.0;
    }
  }
  if (n >= 2) {
    er[n - 2] = ar[(size_t)(n - 1) + (size_t)(n - 2) * n];
    ei[n - 2] = ai[(size_t)(n - 1) + (size_t)(n - 2) * n];
  }
  // real tridiagonal D^+ T D with D[i+1] = D[i] s_i / |s_i|
  std::vector<double> d(n), e(n, 0.0);
  std::vector<cd> D(n, cd(1.0));
  for (int i = 0; i < n; ++i) d[i] = ar[(size_t)i + (size_t)i * n];
  for (int i = 0; i + 1 < n; ++i) {
    const double a = std::hypot(er[i], ei[i]);
    e[i] = a;
    D[i + 1] = a > 0 ? D[i] * cd(er[i] / a, ei[i] / a) : D[i];
  }
  // implicit QL with shifts; rotations recorded (few vectors) or accumulated
  const bool accumulate = nvec > 0 && (size_t)nvec * 4 > (size_t)n;
  std::vector<double> Zc;
  if (accumulate) {
    Zc.assign(nn, 0.0);
    for (int i = 0; i < n; ++i) Zc[(size_t)i * n + i] = 1.0;
  }
  std::vector<int> rot_i;
  std::vector<double> rot_c, rot_s;
  const bool ok = ql_rotations(n, d.data(), e.data(), [&](int i, double c, double s) {
    if (accumulate) {
      double *z0 = &Zc[(size_t)i * n], *z1 = &Zc[(size_t)(i + 1) * n];
      for (int k2 = 0; k2 < n; ++k2) {
        const double f1 = z1[k2];
        z1[k2] = s * z0[k2] + c * f1;
        z0[k2] = c * z0[k2] - s * f1;
      }
    } else if (nvec > 0) {
      rot_i.push_back(i);
      rot_c.push_back(c);
      rot_s.push_back(s);
    }
  });
  if (!ok) return false;
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return d[a] < d[b]; });
  for (int jj = 0; jj < n; ++jj) w[jj] = d[ord[jj]];
  // eigenvectors: tridiagonal vector x, y = D x, then H_{n-3} .. H_0 applied
  std::vector<double> x(n);
  std::vector<double> yr(n), yi(n);
  for (int jj = 0; jj < nvec; ++jj) {
    const int j = ord[jj];
    if (accumulate) {
      std::copy(&Zc[(size_t)j * n], &Zc[(size_t)j * n] + n, x.begin());
    } else {
      // Z = G_1 G_2 ... G_K, so Z e_j applies G_K first
      std::fill(x.begin(), x.end(), 0.0);
      x[j] = 1.0;
      for (size_t t = rot_i.size(); t-- > 0;) {
        const int i = rot_i[t];
        const double c = rot_c[t], s = rot_s[t];
        const double x0 = x[i], x1 = x[i + 1];
        x[i] = c * x0 + s * x1;
        x[i + 1] = -s * x0 + c * x1;
      }
    }
    for (int i = 0; i < n; ++i) {
      yr[i] = D[i].real() * x[i];
      yi[i] = D[i].imag() * x[i];
    }
    for (int k = n - 3; k >= 0; --k) {
      if (bet[k] == 0.0) continue;
      const int m = n - k - 1;
      const double *vr = &ar[(size_t)(k + 1) + (size_t)k * n], *vi = &ai[(size_t)(k + 1) + (size_t)k * n];
      double sr = 0.0, si = 0.0;  // v^+ y
      for (int i = 0; i < m; ++i) {
        sr += vr[i] * yr[k + 1 + i] + vi[i] * yi[k + 1 + i];
        si += vr[i] * yi[k + 1 + i] - vi[i] * yr[k + 1 + i];
      }
      sr *= bet[k];
      si *= bet[k];
      for (int i = 0; i < m; ++i) {
        yr[k + 1 + i] -= vr[i] * sr - vi[i] * si;
        yi[k + 1 + i] -= vr[i] * si + vi[i] * sr;
      }
    }
    for (int i = 0; i < n; ++i) Z[(size_t)i + (size_t)jj * n] = cd(yr[i], yi[i]);
  }
  return true;
}

}  // namespace

bool qs_tridiag_project(int n, const double *d_in, const double *e_in, int nv, const cd *Y, double *w, cd *P) {
  if (n <= 0) return true;
  std::vector<double> d(d_in, d_in + n), e(n, 0.0);
  for (int i = 0; i + 1 < n; ++i) e[i] = e_in[i];
  std::vector<cd> R(Y, Y + (size_t)n * nv);  // row i: component i of every vector
  const bool ok = ql_rotations(n, d.data(), e.data(), [&](int i, double c, double s) {
    cd *r0 = &R[(size_t)i * nv], *r1 = &R[(size_t)(i + 1) * nv];
    for (int k = 0; k < nv; ++k) {
      const cd f1 = r1[k];
      r1[k] = s * r0[k] + c * f1;
      r0[k] = c * r0[k] - s * f1;
    }
  });
  if (!ok) return false;
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return d[a] < d[b]; });
  for (int k = 0; k < n; ++k) {
    w[k] = d[ord[k]];
    std::copy(&R[(size_t)ord[k] * nv], &R[(size_t)ord[k] * nv] + nv, P + (size_t)k * nv);
  }
  return true;
}

bool qs_tridiag_lowest_vectors(int n, const double *d_in, const double *e_in, int k, double *w, double *Z) {
  if (n <= 0) return true;
  k = std::max(0, std::min(k, n));
  std::vector<double> d(d_in, d_in + n), e(n, 0.0);
  for (int i = 0; i + 1 < n; ++i) e[i] = e_in[i];
  std::vector<int> rot_i;
  std::vector<double> rot_c, rot_s;
  const bool ok = ql_rotations(n, d.data(), e.data(), [&](int i, double c, double s) {
    rot_i.push_back(i);
    rot_c.push_back(c);
    rot_s.push_back(s);
  });
  if (!ok) return false;
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return d[a] < d[b]; });
  for (int j = 0; j < n; ++j) w[j] = d[ord[j]];
  for (int j = 0; j < k; ++j) {
    // Z = G_1 G_2 ... G_K, so Z e_j applies G_K first
    double *x = Z + (size_t)j * n;
    std::fill(x, x + n, 0.0);
    x[ord[j]] = 1.0;
    for (size_t t = rot_i.size(); t-- > 0;) {
      const int i = rot_i[t];
      const double c = rot_c[t], s = rot_s[t];
      const double x0 = x[i], x1 = x[i + 1];
      x[i] = c * x0 + s * x1;
      x[i + 1] = -s * x0 + c * x1;
    }
  }
  return true;
}

void qs_heev(int n, const double *A, double *w, double *Z) {
  heev_impl(n, reinterpret_cast<const cd *>(A), w, reinterpret_cast<cd *>(Z), n);
}

bool qs_heev_lowest(int n, const double *A, double *w, double *Z, int nvec) {
  return heev_impl(n, reinterpret_cast<const cd *>(A), w, reinterpret_cast<cd *>(Z), nvec);
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

// Number of eigenvalues below x and S(x) = sum_j 1/(lambda_j - x) = -p'(x)/p(x)
// for p(x) = det(T - x) = prod_i q_i(x), the LDL^T pivots
// q_i = d_i - x - e_{i-1}^2 / q_{i-1}, with q_i' = -1 + e_{i-1}^2 q_{i-1}' / q_{i-1}^2.
int sturm_newton(int m, const double *d, const double *e, double x, double *S) {
  const double tiny = std::numeric_limits<double>::min();
  double q = d[0] - x, dq = -1.0;
  int cnt = q < 0;
  double s = -dq / (q != 0.0 ? q : tiny);
  for (int i = 1; i < m; ++i) {
    const double qq = q != 0.0 ? q : std::numeric_limits<double>::epsilon() * (std::fabs(e[i - 1]) + 1e-300);
    const double t = e[i - 1] * e[i - 1] / qq;
    dq = -1.0 + t * dq / qq;
    q = d[i] - x - t;
    if (q < 0) ++cnt;
    s -= dq / (q != 0.0 ? q : tiny);
  }
  *S = s;
  return cnt;
}

// Lowest eigenvalue by Newton's iteration on det(T - x) from the left of the
// spectrum, where it rises monotonically to lambda_min. A bracket [lo, hi]
// (no eigenvalue below lo, at least one below hi) is kept; Newton runs from
// lo (from hi once the left iterate overshoots, as rounding makes it do next
// to lambda_min), and a bisection step is taken whenever the bracket has not
// halved over the last two steps.
// lower / upper: values expected below / above lambda_min (checked), or NaN.
// Stops once the bracket is narrower than tol (absolute).
double lowest_eigenvalue(int m, const double *d, const double *e, double lower, double upper, double tol) {
  double glo = d[0], ghi = d[0], dmin = d[0];
  for (int i = 0; i < m; ++i) {
    const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(e[i]) : 0.0);
    glo = std::min(glo, d[i] - r);
    ghi = std::max(ghi, d[i] + r);
    dmin = std::min(dmin, d[i]);
  }
  const double span = std::max(ghi - glo, 1e-300);
  double lo = glo - 1e-12 * span, hi = ghi + 1e-12 * span, S_lo = 0.0, S;
  const double h0 = dmin + 1e-12 * span;  // lambda_min <= min d_i
  if (std::isfinite(upper) && upper < h0) {
    if (upper > lo && sturm_newton(m, d, e, upper, &S) > 0) hi = upper;
  }
  if (hi > h0 && h0 < hi && sturm_newton(m, d, e, h0, &S) > 0) hi = h0;
  if (std::isfinite(lower) && lower > lo && lower < hi && sturm_newton(m, d, e, lower, &S) == 0) lo = lower;
  (void)sturm_newton(m, d, e, lo, &S_lo);
  double S_hi = NAN, w1 = INFINITY, w2 = INFINITY;  // bracket widths one and two steps back
  for (int it = 0; it < 400; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (hi - lo <= tol || mid <= lo || mid >= hi) return mid;  // bracket at the tolerance
    double x = mid;
    // Newton unless the bracket failed to halve over the last two steps;
    // steps of at least tol/2, so that a converged iterate closes the bracket
    if (hi - lo <= 0.5 * w2) {
      double cand = NAN;
      if (S_lo > 0 && std::isfinite(S_lo)) cand = lo + std::max(1.0 / S_lo, 0.5 * tol);
      if (!(cand < hi) && S_hi < 0 && std::isfinite(S_hi)) cand = hi + std::min(1.0 / S_hi, -0.5 * tol);
      if (cand > lo && cand < hi) x = cand;
    }
    w2 = w1;
    w1 = hi - lo;
    if (sturm_newton(m, d, e, x, &S) > 0) {
      hi = x;
      S_hi = S;
    } else {
      lo = x;
      S_lo = S;
    }
  }
  return lo;
}

}  // namespace

// Lowest eigenpair of T: eigenvalue by safeguarded Newton on the Sturm
// sequence, eigenvector by inverse iteration with the tridiagonal LU (Thomas
// algorithm, partial pivoting).
static qs_status tridiag_lowest(int m, const double *d, const double *e, double lower, double upper, double *E,
                                double *z) {
  if (m < 1 || !d || !E || !z || (m > 1 && !e)) return QS_EINVAL;
  if (m == 1) {
    *E = d[0];
    z[0] = 1.0;
    return QS_OK;
  }
  double scale = 0.0;
  for (int i = 0; i < m; ++i) scale = std::max(scale, std::fabs(d[i]) + (i + 1 < m ? std::fabs(e[i]) : 0.0));
  scale = std::max(scale, 1e-300);
  const double lam = lowest_eigenvalue(m, d, e, lower, upper, 4 * std::numeric_limits<double>::epsilon() * scale);
  const double shift = lam - 1e-14 * scale;
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
  for (int iter = 0; iter < 3; ++iter) {
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
  *E = lam;
  // fix the sign: largest component positive
  int im = 0;
  for (int i = 1; i < m; ++i)
    if (std::fabs(x[i]) > std::fabs(x[im])) im = i;
  const double sg = x[im] < 0 ? -1.0 : 1.0;
  for (int i = 0; i < m; ++i) z[i] = sg * x[i];
  return QS_OK;
}

extern "C" qs_status qs_tridiag_lowest_host(int m, const double *d, const double *e, double *E, double *z) {
  return tridiag_lowest(m, d, e, std::numeric_limits<double>::quiet_NaN(), std::numeric_limits<double>::quiet_NaN(),
                        E, z);
}

extern "C" qs_status qs_tridiag_lowest_from(int m, const double *d, const double *e, double lower, double upper,
                                            double *E, double *z) {
  return tridiag_lowest(m, d, e, lower, upper, E, z);
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
  if (!heev_impl(n, reinterpret_cast<const cd *>(A), w, reinterpret_cast<cd *>(Z), n)) return QS_ENOTCONVERGED;
  return QS_OK;
}
