// Spectra on the GPU outside the Lanczos recurrence:
//  - the dense small-problem path (the counterpart of the reference's LAPACK
//    zheev): eigenvalues of a dense Hermitian matrix and the projections of
//    the transition vectors on its eigenvectors (Householder reduction on the
//    GPU; the tridiagonal eigenproblem by divide and conquer on the GPU, or by
//    QL on the host for small matrices), the lowest eigenvectors, or all of
//    them, and spectra from the resulting lists of poles
//    (ListOfPolesToSpectrum,
//    Core/Spectral.cpp:1516-1591; BlockListOfPolesToSpectraWithBroadening,
//    Core/BasicMathBlock/BasicMath_BlockTypeConversion.cpp:588-660);
//  - Gaussian and Lorentzian broadening of spectra on an equidistant energy
//    grid (BroadenG / BroadenL / BroadenGvar / BroadenLvar,
//    Core/Spectral.cpp:958-1131).
//
// Broadening:
// out(o) = sum_i in(i) * K(o, i): the kernel K is the source point's line
// shape averaged over 7 sub-points (o - i + j/7, j = -3..3) of the grid cell.
// Grid: one thread per output point o, the source points i split into
// BR_CHUNKS contiguous chunks (blockIdx.y), partial sums per chunk reduced in
// chunk order by k_broaden_sum, so results do not depend on the launch.
// When every source point has the same width, K depends on o - i only and is
// tabulated once (2 npts - 1 values, same arithmetic as the direct form).
#include "qs_internal.h"
#include "qs_kernels.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <vector>

namespace {

constexpr int BR_TILE = 8;     // spectra per thread pass
constexpr int BR_CHUNKS = 16;  // source-point chunks

// Line-shape weight of source point i seen at o - i = d: Gaussian (KIND 0,
// q0 = 2 sigma^2, q1 = dE / (sigma sqrt(2 pi)) / 7) or Lorentzian (KIND 1,
// q0 = FWHM, q1 = 2 dE FWHM / (7 pi)).
template <int KIND>
__device__ __forceinline__ double line_weight(int d, double de, double q0, double q1) {
  double w = 0.0;
#pragma unroll
  for (int j = -3; j <= 3; ++j) {
    const double x = ((double)d + (double)j / 7.0) * de;
    if (KIND == 0)
      w += exp(-(x * x) / q0);
    else
      w += 1.0 / (q0 * q0 + 4.0 * (x * x));
  }
  return w * q1;
}

template <int KIND>
__global__ void k_broaden_table(int npts, double de, double q0, double q1, double *__restrict__ tab) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * npts - 1) return;
  tab[t] = line_weight<KIND>(t - (npts - 1), de, q0, q1);
}

// TABLE: weights from tab[o - i + npts - 1]; else from per-source (q0, q1) in p.
template <int KIND, bool TABLE>
__global__ __launch_bounds__(64) void k_broaden(int nspec, int npts, int chunk, double de,
                                                const double *__restrict__ p, const double *__restrict__ tab,
                                                const double *__restrict__ re, const double *__restrict__ im,
                                                double *__restrict__ part) {
  const int o = blockIdx.x * 64 + threadIdx.x;
  const int c = blockIdx.y;
  const int s0 = blockIdx.z * BR_TILE;
  const int ns = min(BR_TILE, nspec - s0);
  const int i0 = c * chunk, i1 = min(npts, i0 + chunk);
  if (o >= npts) return;
  double ar[BR_TILE], ai[BR_TILE];
#pragma unroll
  for (int s = 0; s < BR_TILE; ++s) ar[s] = ai[s] = 0.0;
  for (int i = i0; i < i1; ++i) {
    const double w = TABLE ? tab[o - i + npts - 1] : line_weight<KIND>(o - i, de, p[2 * i], p[2 * i + 1]);
#pragma unroll
    for (int s = 0; s < BR_TILE; ++s) {
      if (s < ns) {
        ar[s] += re[(size_t)(s0 + s) * npts + i] * w;
        ai[s] += im[(size_t)(s0 + s) * npts + i] * w;
      }
    }
  }
  const size_t plane = (size_t)nspec * npts;
#pragma unroll
  for (int s = 0; s < BR_TILE; ++s) {
    if (s < ns) {
      const size_t at = (size_t)c * 2 * plane + (size_t)(s0 + s) * npts + o;
      part[at] = ar[s];
      part[at + plane] = ai[s];
    }
  }
}

__global__ void k_broaden_sum(int nspec, int npts, int nchunks, const double *__restrict__ part,
                              double *__restrict__ re, double *__restrict__ im) {
  const size_t plane = (size_t)nspec * npts;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= plane) return;
  double sr = 0.0, si = 0.0;
  for (int c = 0; c < nchunks; ++c) {
    sr += part[(size_t)c * 2 * plane + t];
    si += part[(size_t)c * 2 * plane + plane + t];
  }
  re[t] = sr;
  im[t] = si;
}

// (q0, q1) of a Gaussian or Lorentzian of full width f on a grid of step de.
void line_params(int kind, double f, double de, double *q0, double *q1) {
  if (kind == 0) {
    const double sigma = f / (2.0 * std::sqrt(2.0 * std::log(2.0)));
    *q0 = 2.0 * sigma * sigma;
    *q1 = (1.0 / (sigma * std::sqrt(2.0 * M_PI) / de)) / 7.0;
  } else {
    *q0 = f;
    *q1 = de * 2.0 * f / (7.0 * M_PI);
  }
}

}  // namespace

extern "C" qs_status qs_broaden(qs_ctx *ctx, int nspec, int npts, double dE, int kind, int per_point,
                                const double *fwhm, double *re, double *im) {
  if (!ctx || nspec < 1 || npts < 1 || !(dE > 0.0) || (kind != 0 && kind != 1) || !fwhm || !re || !im)
    return QS_EINVAL;
  if (npts > (1 << 24)) return QS_EINVAL;
  for (int i = 0; i < (per_point ? npts : 1); ++i)
    if (std::isnan(fwhm[i])) return QS_EINVAL;
  if (!per_point && 2.0 * fwhm[0] < dE) return QS_OK;  // narrower than the grid: left as it is
  QS_HIP(ctx, hipSetDevice(ctx->device));
  const int nchunks = std::min(BR_CHUNKS, npts);
  const int chunk = (npts + nchunks - 1) / nchunks;
  const size_t plane = (size_t)nspec * npts;
  const size_t np = per_point ? 2 * (size_t)npts : 2 * (size_t)npts - 1;
  char *mem = nullptr;
  QS_TRY(qs_malloc(ctx, (void **)&mem, (2 * plane + 2 * (size_t)nchunks * plane + np) * 8));
  double *dre = (double *)mem, *dim = dre + plane, *part = dim + plane, *dp = part + 2 * (size_t)nchunks * plane;
  std::vector<double> hp;
  double q0 = 0.0, q1 = 0.0;
  if (per_point) {
    hp.resize(2 * (size_t)npts);
    for (int i = 0; i < npts; ++i) line_params(kind, std::max(fwhm[i], dE / 2.0), dE, &hp[2 * i], &hp[2 * i + 1]);
  } else {
    line_params(kind, fwhm[0], dE, &q0, &q1);
  }
  hipError_t e = hipMemcpyAsync(dre, re, plane * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dim, im, plane * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess && per_point) e = hipMemcpyAsync(dp, hp.data(), np * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess && !per_point) {
    const unsigned g = (unsigned)((np + 255) / 256);
    if (kind == 0)
      hipLaunchKernelGGL(k_broaden_table<0>, dim3(g), dim3(256), 0, ctx->stream, npts, dE, q0, q1, dp);
    else
      hipLaunchKernelGGL(k_broaden_table<1>, dim3(g), dim3(256), 0, ctx->stream, npts, dE, q0, q1, dp);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    const dim3 grid((unsigned)((npts + 63) / 64), (unsigned)nchunks, (unsigned)((nspec + BR_TILE - 1) / BR_TILE));
    const double *pp = per_point ? dp : nullptr, *tab = per_point ? nullptr : dp;
    if (kind == 0 && per_point)
      hipLaunchKernelGGL((k_broaden<0, false>), grid, dim3(64), 0, ctx->stream, nspec, npts, chunk, dE, pp, tab, dre,
                         dim, part);
    else if (kind == 0)
      hipLaunchKernelGGL((k_broaden<0, true>), grid, dim3(64), 0, ctx->stream, nspec, npts, chunk, dE, pp, tab, dre,
                         dim, part);
    else if (per_point)
      hipLaunchKernelGGL((k_broaden<1, false>), grid, dim3(64), 0, ctx->stream, nspec, npts, chunk, dE, pp, tab, dre,
                         dim, part);
    else
      hipLaunchKernelGGL((k_broaden<1, true>), grid, dim3(64), 0, ctx->stream, nspec, npts, chunk, dE, pp, tab, dre,
                         dim, part);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_broaden_sum, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, ctx->stream, nspec, npts,
                       nchunks, part, dre, dim);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(re, dre, plane * 8, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(im, dim, plane * 8, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  qs_free(mem);
  if (e != hipSuccess) return qs_hip_fail(ctx, e, "qs_broaden", __FILE__, __LINE__);
  return QS_OK;
}

// ---------------------------------------------------------------------------
// dense small-problem path
// ---------------------------------------------------------------------------
void qs_blas_release(qs_ctx *ctx) {
  if (ctx && ctx->blas) {
    (void)rocblas_destroy_handle((rocblas_handle)ctx->blas);
    ctx->blas = nullptr;
  }
}

// The tridiagonal eigenproblem of the dense method goes to rocSOLVER's divide
// and conquer (dstedc) on the GPU from order ctx->tridiag_dc on (default 256);
// below it the host QL is faster. Measured on the MI355X: dstedc 0.8 / 1.4 /
// 1.9 ms at n = 190 / 800 / 1410 for random and glued-Wilkinson tridiagonal
// matrices (localized eigenvectors, much deflation), about 6 ms at 1410 for
// those of dense Hermitian matrices (little deflation), against 0.5 / 8.7 /
// 23 ms for the host QL with its rotations applied to one vector; dstedc's
// eigenvectors were orthonormal to 3e-15.
// (rocSOLVER's sterf, steqr and stebz + stein took 0.3-50 s at these orders.)

static qs_status blas_handle(qs_ctx *ctx, rocblas_handle *hb) {
  if (!ctx->blas) {
    rocblas_handle b = nullptr;
    if (rocblas_create_handle(&b) != rocblas_status_success)
      return qs_fail(ctx, QS_EHIP, "rocblas_create_handle failed");
    ctx->blas = b;
  }
  *hb = (rocblas_handle)ctx->blas;
  if (rocblas_set_stream(*hb, ctx->stream) != rocblas_status_success)
    return qs_fail(ctx, QS_EHIP, "rocblas_set_stream failed");
  return QS_OK;
}

namespace {

// P[k * nv + s] = z_k^T y_s for real tridiagonal eigenvectors Z (column k at
// Z + k*n); one wave per (k, s), lanes over the rows with a fixed stride and
// a fixed butterfly, so the sums do not depend on the launch.
__global__ __launch_bounds__(64) void k_project_real(int n, const double *__restrict__ Z,
                                                     const double2 *__restrict__ Y, int nv, double2 *__restrict__ P) {
  const int k = blockIdx.x, sv = blockIdx.y;
  const double *z = Z + (size_t)k * n;
  const double2 *y = Y + (size_t)sv * n;
  double re = 0.0, im = 0.0;
  for (int r = threadIdx.x; r < n; r += 64) {
    const double a = z[r];
    const double2 b = y[r];
    re += a * b.x;
    im += a * b.y;
  }
  re = wave_sum(re);
  im = wave_sum(im);
  if (threadIdx.x == 0) P[(size_t)k * nv + sv] = make_double2(re, im);
}

// V[t] = Z[t] + 0i for t < count
__global__ void k_real_to_complex(int64_t count, const double *__restrict__ Z, double2 *__restrict__ V) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) V[t] = make_double2(Z[t], 0.0);
}

// G_s(w_i) = sum_k |P[k][s]|^2 / (w_i + E0_s - E_k + i Gamma/2), w_i = i de + emin.
__global__ void k_poles(int npole, int nv, const double *__restrict__ E, const double2 *__restrict__ P,
                        const double *__restrict__ E0, double emin, double de, int ne, double gamma,
                        double2 *__restrict__ out) {
  const int64_t np = (int64_t)ne + 1;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= np * nv) return;
  const int sv = (int)(idx / np);
  const int i = (int)(idx % np);
  const double om = (double)i * de + emin + E0[sv];
  const double g2 = gamma / 2.0;
  double re = 0.0, im = 0.0;
  for (int k = 0; k < npole; ++k) {
    const double2 p = P[(size_t)k * nv + sv];
    const double b = p.x * p.x + p.y * p.y;
    const double w = om - E[k];
    const double den = g2 * g2 + w * w;
    re += b * w / den;
    im += b * (-g2) / den;
  }
  out[idx] = make_double2(re, im);
}

// G[i][r][c] = sum_k conj(P[k][r]) P[k][c] / (w_i - (E_k - E0) + i Gamma/2),
// w_i = (i / ne) (emax - emin) + emin.
__global__ void k_poles_tensor(int npole, int nv, const double *__restrict__ E, const double2 *__restrict__ P,
                               double E0, double emin, double emax, int ne, double gamma,
                               double2 *__restrict__ out) {
  const int64_t nb = (int64_t)nv * nv;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ((int64_t)ne + 1) * nb) return;
  const int i = (int)(idx / nb);
  const int r = (int)((idx % nb) / nv);
  const int c = (int)(idx % nv);
  const double om = ((double)i / (double)ne) * (emax - emin) + emin;
  const double g2 = gamma / 2.0;
  double re = 0.0, im = 0.0;
  for (int k = 0; k < npole; ++k) {
    const double a = E[k] - E0;
    const double den = g2 * g2 + (om - a) * (om - a);
    const double lr = (om - a) / den, li = (-g2) / den;
    const double2 pc = P[(size_t)k * nv + c], pr = P[(size_t)k * nv + r];
    const double mr = pc.x * pr.x + pc.y * pr.y;
    const double mi = -pc.x * pr.y + pc.y * pr.x;
    re += mr * lr - mi * li;
    im += mr * li + mi * lr;
  }
  out[idx] = make_double2(re, im);
}

}  // namespace

// Y <- Q^+ Y for the Householder reduction of zhetrd (lower): Q = H_0 H_1 ...
// H_{n-2}, H_k = I - tau_k v_k v_k^+, v_k = (0..0, 1, A[k+2..n-1, k]), so
// Q^+ Y applies H_0^+ first. One workgroup per vector, reflectors in order,
// the dot products summed in a fixed order; tau_k = 0 (a column already
// reduced) is skipped. rocSOLVER's zunmtr returns NaN when some tau_k is 0,
// as in block-diagonal (symmetric) Hamiltonians.
__global__ __launch_bounds__(256) void k_apply_qh(int n, const double2 *__restrict__ A,
                                                  const double2 *__restrict__ tau, double2 *__restrict__ Y) {
  __shared__ double red[4];
  double2 *y = Y + (size_t)blockIdx.x * n;
  for (int k = 0; k + 1 < n; ++k) {
    const double2 t = tau[k];
    if (t.x == 0.0 && t.y == 0.0) continue;
    const double2 *col = A + (size_t)k * n;
    double sr = 0.0, si = 0.0;  // v^+ y
    for (int i = k + 1 + threadIdx.x; i < n; i += blockDim.x) {
      const double2 v = i == k + 1 ? make_double2(1.0, 0.0) : col[i];
      const double2 w = y[i];
      sr += v.x * w.x + v.y * w.y;
      si += v.x * w.y - v.y * w.x;
    }
    sr = block_sum_values(sr, red);
    si = block_sum_values(si, red);
    // y -= conj(tau) v (v^+ y)
    const double fr = t.x * sr + t.y * si, fi = t.x * si - t.y * sr;
    for (int i = k + 1 + threadIdx.x; i < n; i += blockDim.x) {
      const double2 v = i == k + 1 ? make_double2(1.0, 0.0) : col[i];
      double2 w = y[i];
      w.x -= v.x * fr - v.y * fi;
      w.y -= v.x * fi + v.y * fr;
      y[i] = w;
    }
    __syncthreads();
  }
}

// Y <- Q Y (the back-transform of tridiagonal eigenvectors): H_{n-2} first,
// y -= tau v (v^+ y). Layout and order of sums as in k_apply_qh. (Blocks of 64
// reflectors as compact WY products applied by rocBLAS zgemm took as long at
// n = 1410, 24.1 against 23.5 ms for all eigenvectors, and longer at 500 and
// 1000: the GEMM shapes, 64-column panels and a 64 x 64 Gram matrix over the
// rows, ran at about 2.4 TFLOP/s.)
__global__ __launch_bounds__(256) void k_apply_q(int n, const double2 *__restrict__ A,
                                                 const double2 *__restrict__ tau, double2 *__restrict__ Y) {
  __shared__ double red[4];
  double2 *y = Y + (size_t)blockIdx.x * n;
  for (int k = n - 2; k >= 0; --k) {
    const double2 t = tau[k];
    if (t.x == 0.0 && t.y == 0.0) continue;
    const double2 *col = A + (size_t)k * n;
    double sr = 0.0, si = 0.0;  // v^+ y
    for (int i = k + 1 + threadIdx.x; i < n; i += blockDim.x) {
      const double2 v = i == k + 1 ? make_double2(1.0, 0.0) : col[i];
      const double2 w = y[i];
      sr += v.x * w.x + v.y * w.y;
      si += v.x * w.y - v.y * w.x;
    }
    sr = block_sum_values(sr, red);
    si = block_sum_values(si, red);
    const double fr = t.x * sr - t.y * si, fi = t.x * si + t.y * sr;  // tau (v^+ y)
    for (int i = k + 1 + threadIdx.x; i < n; i += blockDim.x) {
      const double2 v = i == k + 1 ? make_double2(1.0, 0.0) : col[i];
      double2 w = y[i];
      w.x -= v.x * fr - v.y * fi;
      w.y -= v.x * fi + v.y * fr;
      y[i] = w;
    }
    __syncthreads();
  }
}

#define QS_AQ_RPT 6  // rows per thread of k_apply_q_reg: n <= 1536

// Y <- Q Y as k_apply_q, for n <= 256 QS_AQ_RPT: the workgroup's column in
// registers (thread t holds rows t, t + 256, ...), reflector k-1's rows loaded
// while step k reduces, one barrier per reflector (wave sums by DPP, the four
// wave totals through double-buffered LDS, added in wave order).
__global__ __launch_bounds__(256) void k_apply_q_reg(int n, const double2 *__restrict__ A,
                                                     const double2 *__restrict__ tau, double2 *__restrict__ Y) {
  __shared__ double red[2][8];
  const int tid = (int)threadIdx.x, wv = tid >> 6;
  double2 *y = Y + (size_t)blockIdx.x * n;
  double2 yr[QS_AQ_RPT], vc[QS_AQ_RPT], vx[QS_AQ_RPT];
#pragma unroll
  for (int r = 0; r < QS_AQ_RPT; ++r) {
    const int i = tid + 256 * r;
    yr[r] = i < n ? y[i] : make_double2(0.0, 0.0);
  }
  // rows of reflector k: zero above row k+1, its implicit unit at k+1
  auto load_v = [&](int k, double2(&v)[QS_AQ_RPT]) {
    const double2 *col = A + (size_t)k * n;
#pragma unroll
    for (int r = 0; r < QS_AQ_RPT; ++r) {
      const int i = tid + 256 * r;
      v[r] = i > k + 1 && i < n ? col[i] : make_double2(i == k + 1 ? 1.0 : 0.0, 0.0);
    }
  };
  if (n < 2) return;
  load_v(n - 2, vc);
  int buf = 0;
  for (int k = n - 2; k >= 0; --k) {
    if (k > 0) load_v(k - 1, vx);
    const double2 t = tau[k];
    if (t.x != 0.0 || t.y != 0.0) {
      double sr = 0.0, si = 0.0;  // v^+ y
#pragma unroll
      for (int r = 0; r < QS_AQ_RPT; ++r) {
        sr += vc[r].x * yr[r].x + vc[r].y * yr[r].y;
        si += vc[r].x * yr[r].y - vc[r].y * yr[r].x;
      }
      sr = wave_total(sr);
      si = wave_total(si);
      if ((tid & 63) == 0) {
        red[buf][2 * wv] = sr;
        red[buf][2 * wv + 1] = si;
      }
      __syncthreads();
      sr = (red[buf][0] + red[buf][2]) + (red[buf][4] + red[buf][6]);
      si = (red[buf][1] + red[buf][3]) + (red[buf][5] + red[buf][7]);
      buf ^= 1;
      const double fr = t.x * sr - t.y * si, fi = t.x * si + t.y * sr;  // tau (v^+ y)
#pragma unroll
      for (int r = 0; r < QS_AQ_RPT; ++r) {
        yr[r].x -= vc[r].x * fr - vc[r].y * fi;
        yr[r].y -= vc[r].x * fi + vc[r].y * fr;
      }
    }
    if (k > 0)
#pragma unroll
      for (int r = 0; r < QS_AQ_RPT; ++r) vc[r] = vx[r];
  }
#pragma unroll
  for (int r = 0; r < QS_AQ_RPT; ++r) {
    const int i = tid + 256 * r;
    if (i < n) y[i] = yr[r];
  }
}

// Y <- Q Y for ncols columns of Y (n x ncols, column-major).
static hipError_t launch_apply_q(qs_ctx *ctx, int n, const double2 *A, const double2 *tau, double2 *Y, int ncols) {
  if (n < 2 || ncols < 1) return hipSuccess;
  if (n <= 256 * QS_AQ_RPT)
    hipLaunchKernelGGL(k_apply_q_reg, dim3((unsigned)ncols), dim3(256), 0, ctx->stream, n, A, tau, Y);
  else
    hipLaunchKernelGGL(k_apply_q, dim3((unsigned)ncols), dim3(256), 0, ctx->stream, n, A, tau, Y);
  return hipGetLastError();
}

// Householder reduction A = Q T Q^+ (T real symmetric tridiagonal: d, e; A
// holds the reflectors, tau their factors) with Y <- Q^+ Y for nv vectors: the
// library's cooperative reduction when it takes the order, rocSOLVER's zhetrd
// and k_apply_qh otherwise.
static qs_status reduce_dense(qs_ctx *ctx, rocblas_handle hb, int n, double *A, double *dd, double *de,
                              double2 *dtau, int nv, double2 *dY) {
  if (ctx->hetrd) {
    const qs_status st = qs_hetrd_fused(ctx, n, (double2 *)A, dd, de, dtau, nv, dY);
    if (st != QS_EUNSUPPORTED) return st;
  }
  const rocblas_status rs = rocsolver_zhetrd(hb, rocblas_fill_lower, n, (rocblas_double_complex *)A, n, dd, de,
                                             (rocblas_double_complex *)dtau);
  if (rs != rocblas_status_success) return qs_fail(ctx, QS_EHIP, "rocsolver_zhetrd failed (status %d)", (int)rs);
  if (nv > 0) {
    hipLaunchKernelGGL(k_apply_qh, dim3((unsigned)nv), dim3(256), 0, ctx->stream, n, (const double2 *)A,
                       (const double2 *)dtau, dY);
    QS_HIP(ctx, hipGetLastError());
  }
  return QS_OK;
}

// Eigenvalues (dd, ascending; de destroyed) and eigenvectors Z (n x n real,
// column k = eigenvector k) of the tridiagonal T = (dd, de) on the GPU by
// rocSOLVER's divide and conquer. info: one device int.
static qs_status tridiag_dc(qs_ctx *ctx, rocblas_handle hb, int n, double *dd, double *de, double *Z,
                            rocblas_int *info) {
  const rocblas_status rs = rocsolver_dstedc(hb, rocblas_evect_tridiagonal, n, dd, de, Z, n, info);
  if (rs != rocblas_status_success) return qs_fail(ctx, QS_EHIP, "rocsolver_dstedc failed (status %d)", (int)rs);
  rocblas_int hinfo = 0;
  QS_TRY(qs_d2h(ctx, &hinfo, info, sizeof(hinfo)));
  if (hinfo != 0) return qs_fail(ctx, QS_ENOTCONVERGED, "the tridiagonal divide and conquer failed (info %d)", hinfo);
  return QS_OK;
}

// Eigenvalues and projections: the reduction with Y = Q^+ X on the GPU; then
// either (n >= ctx->tridiag_dc, or all eigenvectors wanted) the tridiagonal
// eigenvectors Z on the GPU and P = Z^T Y there, the eigenvectors of A being
// Q Z (k_apply_q, one workgroup per column); or the tridiagonal QL on the host
// with its rotations applied to the rows of Y, which turns Y into Z^T Y.
static qs_status dense_poles_householder(qs_ctx *ctx, rocblas_handle hb, int n, double *A, int nv, const double *X,
                                         double *E, double *P, bool vectors) {
  const bool dc = vectors || n >= ctx->tridiag_dc;
  const size_t ybytes = (size_t)n * std::max(nv, 1) * 16;
  const size_t zbytes = dc ? (size_t)n * n * 8 : 0;
  const size_t vbytes = vectors ? (size_t)n * n * 16 : 0;
  const size_t pbytes = dc ? ybytes : 0;
  char *mem = nullptr;
  // 16-byte arrays first, then the 8-byte ones and the info word
  QS_TRY(qs_malloc(ctx, (void **)&mem, ybytes + pbytes + vbytes + (size_t)n * 16 + zbytes + (size_t)n * 16 + 256));
  double2 *dY = (double2 *)mem;
  double2 *dP = (double2 *)(mem + ybytes);
  double2 *dV = (double2 *)(mem + ybytes + pbytes);
  double2 *dtau = (double2 *)(mem + ybytes + pbytes + vbytes);
  double *dZ = (double *)(dtau + n);
  double *dd = (double *)((char *)dZ + zbytes), *de = dd + n;
  rocblas_int *info = (rocblas_int *)(de + n);
  qs_status st = QS_OK;
  hipError_t e = nv ? hipMemcpyAsync(dY, X, (size_t)n * nv * 16, hipMemcpyDeviceToDevice, ctx->stream) : hipSuccess;
  if (e == hipSuccess) st = reduce_dense(ctx, hb, n, A, dd, de, dtau, nv, dY);
  if (e == hipSuccess && st == QS_OK && dc) {
    st = tridiag_dc(ctx, hb, n, dd, de, dZ, info);
    if (st == QS_OK && nv > 0) {
      hipLaunchKernelGGL(k_project_real, dim3((unsigned)n, (unsigned)nv), dim3(64), 0, ctx->stream, n,
                         (const double *)dZ, (const double2 *)dY, nv, dP);
      e = hipGetLastError();
    }
    if (st == QS_OK && e == hipSuccess && vectors) {
      const int64_t cnt = (int64_t)n * n;
      hipLaunchKernelGGL(k_real_to_complex, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, ctx->stream, cnt,
                         (const double *)dZ, dV);
      e = hipGetLastError();
      if (e == hipSuccess) e = launch_apply_q(ctx, n, (const double2 *)A, (const double2 *)dtau, dV, n);
      if (e == hipSuccess) e = hipMemcpyAsync(A, dV, (size_t)n * n * 16, hipMemcpyDeviceToDevice, ctx->stream);
    }
    if (st == QS_OK && e == hipSuccess) e = hipMemcpyAsync(E, dd, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (st == QS_OK && e == hipSuccess && nv > 0)
      e = hipMemcpyAsync(P, dP, (size_t)n * nv * 16, hipMemcpyDeviceToHost, ctx->stream);
    if (st == QS_OK && e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    else (void)hipStreamSynchronize(ctx->stream);
    qs_free(mem);
    if (st != QS_OK) return st;
    if (e != hipSuccess) return qs_hip_fail(ctx, e, "qs_dense_poles", __FILE__, __LINE__);
    return QS_OK;
  }
  std::vector<double> hd(n), he(n);
  std::vector<std::complex<double>> hy((size_t)n * nv);
  if (e == hipSuccess && st == QS_OK) e = hipMemcpyAsync(hd.data(), dd, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess && st == QS_OK && n > 1)
    e = hipMemcpyAsync(he.data(), de, (size_t)(n - 1) * 8, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess && st == QS_OK && nv > 0)
    e = hipMemcpyAsync(hy.data(), dY, (size_t)n * nv * 16, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess && st == QS_OK) e = hipStreamSynchronize(ctx->stream);
  else (void)hipStreamSynchronize(ctx->stream);
  qs_free(mem);
  if (st != QS_OK) return st;
  if (e != hipSuccess) return qs_hip_fail(ctx, e, "qs_dense_poles", __FILE__, __LINE__);
  // rows of Y: component i of every vector (hy holds vector s at s*n)
  std::vector<std::complex<double>> rows((size_t)n * nv), proj((size_t)n * nv);
  for (int sv = 0; sv < nv; ++sv)
    for (int i = 0; i < n; ++i) rows[(size_t)i * nv + sv] = hy[(size_t)sv * n + i];
  if (!qs_tridiag_project(n, hd.data(), he.data(), nv, rows.data(), E, proj.data()))
    return qs_fail(ctx, QS_ENOTCONVERGED, "the tridiagonal QL did not converge");
  std::copy(proj.begin(), proj.end(), reinterpret_cast<std::complex<double> *>(P));
  return QS_OK;
}

extern "C" qs_status qs_dense_poles(qs_ctx *ctx, int n, double *A, int nv, const double *X, double *E, double *P,
                                    int vectors) {
  if (!ctx || n < 1 || !A || nv < 0 || !E || (nv > 0 && (!X || !P))) return QS_EINVAL;
  QS_HIP(ctx, hipSetDevice(ctx->device));
  rocblas_handle hb = nullptr;
  QS_TRY(blas_handle(ctx, &hb));
  return dense_poles_householder(ctx, hb, n, A, nv, X, E, P, vectors != 0);
}

extern "C" qs_status qs_dense_lowest(qs_ctx *ctx, int n, double *A, int k, double *E, double *V) {
  if (!ctx || n < 1 || !A || k < 0 || k > n || !E || (k > 0 && !V)) return QS_EINVAL;
  QS_HIP(ctx, hipSetDevice(ctx->device));
  rocblas_handle hb = nullptr;
  QS_TRY(blas_handle(ctx, &hb));
  const bool dc = n >= ctx->tridiag_dc;
  const size_t zbytes = dc ? (size_t)n * n * 8 : 0;
  char *mem = nullptr;
  QS_TRY(qs_malloc(ctx, (void **)&mem, (size_t)n * 32 + zbytes + 256));
  double2 *dtau = (double2 *)mem;
  double *dd = (double *)(mem + (size_t)n * 16), *de = dd + n;
  double *dZ = de + n;
  rocblas_int *info = (rocblas_int *)((char *)dZ + zbytes);
  qs_status st = reduce_dense(ctx, hb, n, A, dd, de, dtau, 0, nullptr);
  hipError_t e = hipSuccess;
  std::vector<double> z;
  if (st == QS_OK && dc) {
    // eigenvectors of T on the GPU; the k lowest become V's columns
    st = tridiag_dc(ctx, hb, n, dd, de, dZ, info);
    if (st == QS_OK) e = hipMemcpyAsync(E, dd, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (st == QS_OK && e == hipSuccess && k > 0) {
      const int64_t cnt = (int64_t)n * k;
      hipLaunchKernelGGL(k_real_to_complex, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, ctx->stream, cnt,
                         (const double *)dZ, (double2 *)V);
      e = hipGetLastError();
    }
  } else if (st == QS_OK) {
    std::vector<double> hd(n), he(n);
    e = hipMemcpyAsync(hd.data(), dd, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && n > 1)
      e = hipMemcpyAsync(he.data(), de, (size_t)(n - 1) * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    z.resize((size_t)n * k);
    if (e == hipSuccess && !qs_tridiag_lowest_vectors(n, hd.data(), he.data(), k, E, z.data()))
      st = qs_fail(ctx, QS_ENOTCONVERGED, "the tridiagonal QL did not converge");
    if (st == QS_OK && e == hipSuccess && k > 0) {
      std::vector<double2> zc((size_t)n * k);
      for (size_t t = 0; t < zc.size(); ++t) zc[t] = make_double2(z[t], 0.0);
      e = hipMemcpyAsync(V, zc.data(), zc.size() * 16, hipMemcpyHostToDevice, ctx->stream);
      // the staging vector must outlive the copy
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    }
  }
  if (st == QS_OK && e == hipSuccess && k > 0 && n > 1)
    e = launch_apply_q(ctx, n, (const double2 *)A, (const double2 *)dtau, (double2 *)V, k);
  if (st == QS_OK && e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  else (void)hipStreamSynchronize(ctx->stream);
  qs_free(mem);
  if (st != QS_OK) return st;
  if (e != hipSuccess) return qs_hip_fail(ctx, e, "qs_dense_lowest", __FILE__, __LINE__);
  return QS_OK;
}

extern "C" qs_status qs_pole_spectrum(qs_ctx *ctx, int npole, int nv, const double *E, const double *P, int tensor,
                                      const double *E0, double emin, double emax, int ne, double gamma, double *out) {
  if (!ctx || npole < 1 || nv < 1 || ne < 1 || !E || !P || !E0 || !out) return QS_EINVAL;
  QS_HIP(ctx, hipSetDevice(ctx->device));
  const size_t nout = tensor ? ((size_t)ne + 1) * nv * nv : ((size_t)ne + 1) * nv;
  char *mem = nullptr;
  QS_TRY(qs_malloc(ctx, (void **)&mem, (size_t)npole * 8 + (size_t)npole * nv * 16 + (size_t)nv * 8 + nout * 16));
  double2 *dP = (double2 *)mem;
  double2 *dout = dP + (size_t)npole * nv;
  double *dE = (double *)(dout + nout);
  double *dE0 = dE + npole;
  hipError_t e = hipMemcpyAsync(dE, E, (size_t)npole * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dP, P, (size_t)npole * nv * 16, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dE0, E0, (size_t)nv * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    const unsigned g = (unsigned)((nout + 255) / 256);
    if (tensor)
      hipLaunchKernelGGL(k_poles_tensor, dim3(g), dim3(256), 0, ctx->stream, npole, nv, dE, dP, E0[0], emin, emax, ne,
                         gamma, dout);
    else
      hipLaunchKernelGGL(k_poles, dim3(g), dim3(256), 0, ctx->stream, npole, nv, dE, dP, dE0, emin,
                         (emax - emin) / (double)ne, ne, gamma, dout);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, nout * 16, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  qs_free(mem);
  if (e != hipSuccess) return qs_hip_fail(ctx, e, "qs_pole_spectrum", __FILE__, __LINE__);
  return QS_OK;
}
