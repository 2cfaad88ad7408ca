// Block algorithms on the device: block Lanczos for tensor spectra
// (LanczosBlockTriDiagonalizeCC, Core/BlockLanczos.cpp:2131-2371), the tensor
// continued fraction (BlockTriDiagonalMatrixToSpectraWithBroadening,
// BasicMath_BlockTypeConversion.cpp:1636-1726), and the lowest eigenstates by
// thick-restarted block Lanczos with full re-orthogonalization (the role of
// BlockGroundState, Core/BlockLanczos.cpp:5782-5959).
//
// A block of k <= 16 vectors is stored interleaved, entry j of vector v at
// [j*ld + v], so one gather per nonzero of the assembled operator fetches all
// k values of a column. Block inner products are reduced in a fixed order
// (per-block partials, then one block per output element), so results are
// reproducible run to run.
#include <algorithm>
#include <cmath>
#include <complex>
#include <numeric>
#include <cstring>
#include <vector>

#include "qs_internal.h"
#include "qs_kernels.h"

using cd = std::complex<double>;

#define QS_KMAX 16
#define QS_BDOT_BLOCKS 512

qs_status qs_hop_apply_strided(qs_ctx *ctx, const qs_hop *h, int k, const double *x, int64_t ldx, double *y,
                               int64_t ldy);
qs_status qs_launch_mf_ext(qs_ctx *ctx, const qs_hop *h, const double *x, double *y);

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// uniform in [-1, 1) for re and im, counter-based on (seed, index)
__global__ void k_rand(int64_t n, uint64_t seed, double2 *__restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t a = splitmix64(seed ^ (0xD1B54A32D192ED03ULL * (uint64_t)(2 * i + 1)));
    const uint64_t b = splitmix64(a ^ 0x8BB84B93962EACC9ULL);
    out[i] = make_double2((double)(a >> 11) * (2.0 / 9007199254740992.0) - 1.0,
                          (double)(b >> 11) * (2.0 / 9007199254740992.0) - 1.0);
  }
}

// vector-major (k vectors of n) <-> interleaved [n][k]
__global__ void k_pack(int64_t n, int k, const double2 *__restrict__ src, double2 *__restrict__ dst, int to_interleaved) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n * k; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / k;
    const int v = (int)(i % k);
    if (to_interleaved) dst[i] = src[(int64_t)v * n + r];
    else dst[(int64_t)v * n + r] = src[i];
  }
}

// column v of an interleaved block <-> contiguous vector
__global__ void k_column(int64_t n, const double2 *__restrict__ src, int64_t lds, double2 *__restrict__ dst,
                         int64_t ldd) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
    dst[r * ldd] = src[r * lds];
}

// partial block inner products C[i][j] = sum_r conj(X[r][i]) Y[r][j]
__global__ void __launch_bounds__(QS_BLOCK)
k_bdot(int64_t n, int kx, int ky, const double2 *__restrict__ X, int64_t ldx, const double2 *__restrict__ Y,
       int64_t ldy, double2 *__restrict__ part) {
  __shared__ double2 red[QS_BLOCK];
  const int np = kx * ky;
  const int groups = blockDim.x / np;
  const int t = threadIdx.x;
  const int pair = t % np, grp = t / np;
  double sr = 0.0, si = 0.0;
  if (grp < groups) {
    const int i = pair / ky, j = pair % ky;
    for (int64_t r = (int64_t)blockIdx.x * groups + grp; r < n; r += (int64_t)gridDim.x * groups) {
      const double2 a = X[r * ldx + i], b = Y[r * ldy + j];
      sr += a.x * b.x + a.y * b.y;
      si += a.x * b.y - a.y * b.x;
    }
  }
  red[t] = make_double2(sr, si);
  __syncthreads();
  if (t < np) {
    double2 s = red[t];
    for (int g = 1; g < groups; ++g) {
      s.x += red[g * np + t].x;
      s.y += red[g * np + t].y;
    }
    part[(size_t)blockIdx.x * np + t] = s;
  }
}

// one block per output element: fixed-order sum of the per-block partials
__global__ void __launch_bounds__(QS_BLOCK)
k_bdot_reduce(int nparts, int np, const double2 *__restrict__ part, double2 *__restrict__ out) {
  __shared__ double red[QS_BLOCK];
  const int e = blockIdx.x;
  double sr = 0.0, si = 0.0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
    sr += part[(size_t)b * np + e].x;
    si += part[(size_t)b * np + e].y;
  }
  const double r = block_sum_values(sr, red);
  const double i = block_sum_values(si, red);
  if (threadIdx.x == 0) out[e] = make_double2(r, i);
}

// Y = beta*Y + X C   (X: n x kx, C: kx x ky row-major, Y: n x ky)
__global__ void __launch_bounds__(QS_BLOCK)
k_bupd(int64_t n, int kx, int ky, const double2 *__restrict__ X, int64_t ldx, const double2 *__restrict__ C,
       double2 *__restrict__ Y, int64_t ldy, double beta) {
  __shared__ double2 sc[QS_KMAX * QS_KMAX];
  for (int i = threadIdx.x; i < kx * ky; i += blockDim.x) sc[i] = C[i];
  __syncthreads();
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    double2 x[QS_KMAX];
#pragma unroll
    for (int i = 0; i < QS_KMAX; ++i)
      if (i < kx) x[i] = X[r * ldx + i];
    for (int j = 0; j < ky; ++j) {
      double ar = 0.0, ai = 0.0;
      if (beta != 0.0) {
        const double2 y = Y[r * ldy + j];
        ar = beta * y.x;
        ai = beta * y.y;
      }
#pragma unroll
      for (int i = 0; i < QS_KMAX; ++i) {
        if (i < kx) {
          const double2 c = sc[i * ky + j];
          ar = fma(x[i].x, c.x, fma(-x[i].y, c.y, ar));
          ai = fma(x[i].x, c.y, fma(x[i].y, c.x, ai));
        }
      }
      Y[r * ldy + j] = make_double2(ar, ai);
    }
  }
}

// Tensor continued fraction: one workgroup per energy point. LDS holds the
// running p x p block, a product buffer and the Gauss-Jordan work matrix.
__global__ void __launch_bounds__(QS_BLOCK)
k_block_cf(int k, int n, const double2 *__restrict__ A, const double2 *__restrict__ B, const int *__restrict__ kept,
           double E0, double emin, double de, double gamma, double2 *__restrict__ out) {
  constexpr int L2 = 2 * QS_KMAX;
  __shared__ double2 G[QS_KMAX * QS_KMAX];
  __shared__ double2 T[QS_KMAX * QS_KMAX];
  __shared__ double2 W[QS_KMAX * L2];
  __shared__ double2 fcol[QS_KMAX];
  __shared__ int piv;
  const int t = threadIdx.x;
  const int nt = blockDim.x;
  const int iw = blockIdx.x;
  const double w = emin + iw * de + E0;
  const int kk = k * k;
  for (int lvl = n; lvl >= 1; --lvl) {
    const int p = kept[lvl];
    const int q = lvl < n ? kept[lvl + 1] : 0;
    const double2 *Al = A + (size_t)lvl * kk;
    const double2 *Bl = B + (size_t)lvl * kk;  // B_lvl = V_lvl^+ H V_lvl+1 (p x q)
    __syncthreads();
    for (int e = t; e < p * q; e += nt) {  // T = B_l G_{l+1}
      const int i = e / q, j = e % q;
      double sr = 0.0, si = 0.0;
      for (int m = 0; m < q; ++m) {
        const double2 b = Bl[i * k + m], g = G[m * QS_KMAX + j];
        sr += b.x * g.x - b.y * g.y;
        si += b.x * g.y + b.y * g.x;
      }
      T[i * QS_KMAX + j] = make_double2(sr, si);
    }
    __syncthreads();
    for (int e = t; e < p * p; e += nt) {  // W = [ z - A_l - T B_l^+ | I ]
      const int i = e / p, j = e % p;
      double sr = 0.0, si = 0.0;
      for (int m = 0; m < q; ++m) {
        const double2 a = T[i * QS_KMAX + m], b = Bl[j * k + m];
        sr += a.x * b.x + a.y * b.y;
        si += a.y * b.x - a.x * b.y;
      }
      const double2 av = Al[i * k + j];
      double2 v = make_double2(-av.x - sr, -av.y - si);
      if (i == j) {
        v.x += w;
        v.y += 0.5 * gamma;
      }
      W[i * L2 + j] = v;
      W[i * L2 + QS_KMAX + j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    }
    __syncthreads();
    for (int c = 0; c < p; ++c) {  // Gauss-Jordan, partial pivoting
      if (t == 0) {
        int best = c;
        double bm = -1.0;
        for (int r = c; r < p; ++r) {
          const double2 v = W[r * L2 + c];
          const double m2 = v.x * v.x + v.y * v.y;
          if (m2 > bm) {
            bm = m2;
            best = r;
          }
        }
        piv = best;
      }
      __syncthreads();
      const int pr = piv;
      if (pr != c) {
        for (int e = t; e < L2; e += nt) {
          const double2 tmp = W[c * L2 + e];
          W[c * L2 + e] = W[pr * L2 + e];
          W[pr * L2 + e] = tmp;
        }
      }
      __syncthreads();
      const double2 d = W[c * L2 + c];
      const double dn = d.x * d.x + d.y * d.y;
      const double2 inv = make_double2(d.x / dn, -d.y / dn);
      __syncthreads();
      for (int e = t; e < L2; e += nt) {
        const double2 v = W[c * L2 + e];
        W[c * L2 + e] = make_double2(v.x * inv.x - v.y * inv.y, v.x * inv.y + v.y * inv.x);
      }
      for (int e = t; e < p; e += nt) fcol[e] = W[e * L2 + c];
      __syncthreads();
      for (int e = t; e < p * L2; e += nt) {
        const int r = e / L2, cc = e % L2;
        if (r == c) continue;
        const double2 f = fcol[r];
        const double2 pv = W[c * L2 + cc];
        const double2 v = W[r * L2 + cc];
        W[r * L2 + cc] = make_double2(v.x - (f.x * pv.x - f.y * pv.y), v.y - (f.x * pv.y + f.y * pv.x));
      }
      __syncthreads();
    }
    for (int e = t; e < p * p; e += nt) {
      const int i = e / p, j = e % p;
      G[i * QS_KMAX + j] = W[i * L2 + QS_KMAX + j];
    }
  }
  __syncthreads();
  // out = B0^* G_1 B0^T   (B0: k x p1)
  const int p1 = kept[1];
  for (int e = t; e < k * p1; e += nt) {  // T[r][j] = sum_m conj(B0[r][m]) G[m][j]
    const int r = e / p1, j = e % p1;
    double sr = 0.0, si = 0.0;
    for (int m = 0; m < p1; ++m) {
      const double2 b = B[r * k + m], g = G[m * QS_KMAX + j];
      sr += b.x * g.x + b.y * g.y;
      si += b.x * g.y - b.y * g.x;
    }
    T[r * QS_KMAX + j] = make_double2(sr, si);
  }
  __syncthreads();
  for (int e = t; e < kk; e += nt) {  // sum_j T[r][j] B0[c][j]
    const int r = e / k, c = e % k;
    double sr = 0.0, si = 0.0;
    for (int j = 0; j < p1; ++j) {
      const double2 a = T[r * QS_KMAX + j], b = B[c * k + j];
      sr += a.x * b.x - a.y * b.y;
      si += a.x * b.y + a.y * b.x;
    }
    out[(size_t)iw * kk + e] = make_double2(sr, si);
  }
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
namespace {

int grid_for(qs_ctx *ctx, int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + QS_BLOCK - 1) / QS_BLOCK, (int64_t)ctx->num_cu * 8));
}

struct Work {
  qs_ctx *ctx;
  int64_t n;
  double2 *part = nullptr;  // QS_BDOT_BLOCKS * 256
  double2 *dC = nullptr;    // 256 complex results / coefficients
  qs_status init(qs_ctx *c, int64_t nn) {
    ctx = c;
    n = nn;
    QS_TRY(qs_malloc(ctx, (void **)&part, (size_t)QS_BDOT_BLOCKS * 256 * 16));
    QS_TRY(qs_malloc(ctx, (void **)&dC, 512 * 16));
    return QS_OK;
  }
  ~Work() {
    if (part) (void)hipFree(part);
    if (dC) (void)hipFree(dC);
  }
  // C[i*ky + j] = <X_i|Y_j>
  qs_status bdot(const double2 *X, int kx, int64_t ldx, const double2 *Y, int ky, int64_t ldy, std::vector<cd> &C) {
    C.assign((size_t)kx * ky, cd(0));
    if (kx == 0 || ky == 0) return QS_OK;
    const int np = kx * ky;
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(QS_BDOT_BLOCKS, (n + 63) / 64));
    hipLaunchKernelGGL(k_bdot, dim3(nb), dim3(QS_BLOCK), 0, ctx->stream, n, kx, ky, X, ldx, Y, ldy, part);
    hipLaunchKernelGGL(k_bdot_reduce, dim3(np), dim3(QS_BLOCK), 0, ctx->stream, nb, np, part, dC);
    QS_HIP(ctx, hipGetLastError());
    QS_HIP(ctx, hipMemcpyAsync(C.data(), dC, (size_t)np * 16, hipMemcpyDeviceToHost, ctx->stream));
    QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return QS_OK;
  }
  // Y = beta*Y + X C
  qs_status bupd(const double2 *X, int kx, int64_t ldx, const std::vector<cd> &C, double2 *Y, int ky, int64_t ldy,
                 double beta) {
    if (ky == 0) return QS_OK;
    if (kx == 0 && beta == 1.0) return QS_OK;
    // the coefficient buffer is reused by the next call: wait for prior kernels
    QS_HIP(ctx, hipMemcpyAsync(dC, C.data(), (size_t)kx * ky * 16, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_bupd, dim3(grid_for(ctx, n)), dim3(QS_BLOCK), 0, ctx->stream, n, kx, ky, X, ldx,
                       (const double2 *)dC, Y, ldy, beta);
    QS_HIP(ctx, hipGetLastError());
    QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return QS_OK;
  }
};

// y (interleaved, k vectors) = H x
qs_status block_apply(qs_ctx *ctx, const qs_hop *h, const double2 *x, int k, int64_t ld, double2 *y, double2 *tmp) {
  if (k == 0) return QS_OK;
  if (h->assembled) return qs_hop_apply_strided(ctx, h, k, (const double *)x, ld, (double *)y, ld);
  const int64_t n = h->rows->n;
  for (int v = 0; v < k; ++v) {
    hipLaunchKernelGGL(k_column, dim3(grid_for(ctx, n)), dim3(QS_BLOCK), 0, ctx->stream, n, x + v, ld, tmp, 1);
    QS_TRY(qs_launch_mf_ext(ctx, h, (const double *)tmp, (double *)(tmp + n)));
    hipLaunchKernelGGL(k_column, dim3(grid_for(ctx, n)), dim3(QS_BLOCK), 0, ctx->stream, n, tmp + n, 1, y + v, ld);
  }
  QS_HIP(ctx, hipGetLastError());
  return QS_OK;
}

// Orthonormalize the k columns of W (in place into V): overlap S = W^+ W,
// S = U diag(l) U^+, keep directions with sqrt(l) > max(sv_cut, 3 sqrt(eps) s_max)
// (the singular-value cut of ComplexWaveFunctionListSVDOrthogonalizeSameBasisOnce,
// Core/ComplexWaveFunctionArithmetic.cpp:1356-1550). Without reduction the
// symmetric transform S^-1/2 is used, otherwise U_keep diag(l^-1/2).
qs_status lowdin(Work &wk, const double2 *W, int k, double2 *V, double sv_cut, int *kept) {
  std::vector<cd> S;
  QS_TRY(wk.bdot(W, k, k, W, k, k, S));
  // S[i*k+j] = <W_i|W_j>, column-major for heev: element (i,j) at i + j*k -> transpose = conj
  std::vector<cd> A((size_t)k * k);
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) A[i + (size_t)j * k] = S[(size_t)i * k + j];
  std::vector<double> l(k);
  std::vector<cd> U((size_t)k * k);
  qs_heev(k, (const double *)A.data(), l.data(), (double *)U.data());
  double lmax = 0.0;
  for (int i = 0; i < k; ++i) lmax = std::max(lmax, l[i]);
  const double smax = std::sqrt(std::max(lmax, 0.0));
  const double cut = std::max(sv_cut, 3.0 * std::sqrt(2.220446049250313e-16) * smax);
  std::vector<int> keep;
  for (int i = k - 1; i >= 0; --i)
    if (l[i] > 0 && std::sqrt(l[i]) > cut) keep.push_back(i);
  const int kp = (int)keep.size();
  *kept = kp;
  if (kp == 0) return QS_OK;
  std::vector<cd> Tm((size_t)k * kp, cd(0));  // row-major k x kp
  if (kp == k) {
    // S^-1/2 = U diag(l^-1/2) U^+
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) {
        cd s = 0;
        for (int m = 0; m < k; ++m) s += U[i + (size_t)m * k] * (1.0 / std::sqrt(l[m])) * std::conj(U[j + (size_t)m * k]);
        Tm[(size_t)i * k + j] = s;
      }
  } else {
    for (int i = 0; i < k; ++i)
      for (int jj = 0; jj < kp; ++jj) {
        const int m = keep[jj];
        Tm[(size_t)i * kp + jj] = U[i + (size_t)m * k] / std::sqrt(l[m]);
      }
  }
  QS_TRY(wk.bupd(W, k, k, Tm, V, kp, kp, 0.0));
  return QS_OK;
}

// W -= X (X^+ W)
qs_status project_out(Work &wk, const double2 *X, int kx, double2 *W, int kw, std::vector<cd> *accum) {
  if (kx == 0 || kw == 0) return QS_OK;
  std::vector<cd> C;
  QS_TRY(wk.bdot(X, kx, kx, W, kw, kw, C));
  if (accum)
    for (size_t i = 0; i < C.size(); ++i) (*accum)[i] += C[i];
  for (auto &c : C) c = -c;
  QS_TRY(wk.bupd(X, kx, kx, C, W, kw, kw, 1.0));
  return QS_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// block Lanczos (tensor spectra)
// ---------------------------------------------------------------------------
extern "C" qs_status qs_block_lanczos(qs_ctx *ctx, const qs_hop *h, int k, const double *V0, int nsteps,
                                      double sv_cut, double *A_out, double *B_out, int *kept_out) {
  if (!ctx || !h || !V0 || !A_out || !B_out || k < 1 || k > QS_KMAX || nsteps < 1) return QS_EINVAL;
  if (h->rows != h->cols) return qs_fail(ctx, QS_EINVAL, "qs_block_lanczos needs rows == cols");
  QS_HIP(ctx, hipSetDevice(ctx->device));
  const int64_t n = h->rows->n;
  const size_t bb = (size_t)n * k * 16;
  char *mem = nullptr;
  QS_TRY(qs_malloc(ctx, (void **)&mem, 5 * bb + (size_t)n * 32 + 256));
  double2 *X0 = (double2 *)mem, *B1 = (double2 *)(mem + bb), *B2 = (double2 *)(mem + 2 * bb);
  double2 *B3 = (double2 *)(mem + 3 * bb), *Wb = (double2 *)(mem + 4 * bb), *tmp = (double2 *)(mem + 5 * bb);
  Work wk;
  qs_status st = wk.init(ctx, n);
  const size_t kk = (size_t)k * k;
  std::vector<cd> Aall((size_t)(nsteps + 1) * kk, cd(0)), Ball((size_t)nsteps * kk, cd(0));
  std::vector<int> kept(nsteps + 1, 0);
  kept[0] = k;
  auto run = [&]() -> qs_status {
    QS_TRY(st);
    hipLaunchKernelGGL(k_pack, dim3(grid_for(ctx, n * k)), dim3(QS_BLOCK), 0, ctx->stream, n, k, (const double2 *)V0, X0, 1);
    QS_HIP(ctx, hipGetLastError());
    int p = 0;
    QS_TRY(lowdin(wk, X0, k, B1, sv_cut, &p));
    kept[1] = p;
    if (p == 0) return QS_OK;
    {  // B0[i][j] = <V1_j|X0_i>
      std::vector<cd> C;
      QS_TRY(wk.bdot(B1, p, p, X0, k, k, C));
      for (int i = 0; i < k; ++i)
        for (int j = 0; j < p; ++j) Ball[(size_t)i * k + j] = C[(size_t)j * k + i];
    }
    double2 *bufs[3] = {B1, B2, B3};
    int icur = 0, inext = 1, iprev = -1;
    double2 *Vprev = nullptr, *Vcur = B1, *Vnext = B2;
    int pprev = 0;
    for (int s = 1; s <= nsteps; ++s) {
      QS_TRY(block_apply(ctx, h, Vcur, p, p, Wb, tmp));
      std::vector<cd> C;
      QS_TRY(wk.bdot(Vcur, p, p, Wb, p, p, C));  // A_s
      for (int i = 0; i < p; ++i)
        for (int j = 0; j < p; ++j) Aall[(size_t)s * kk + (size_t)i * k + j] = C[(size_t)i * p + j];
      if (s >= 2) {
        QS_TRY(wk.bdot(Vprev, pprev, pprev, Wb, p, p, C));  // B_{s-1} = V_{s-1}^+ H V_s
        for (int i = 0; i < pprev; ++i)
          for (int j = 0; j < p; ++j) Ball[(size_t)(s - 1) * kk + (size_t)i * k + j] = C[(size_t)i * p + j];
      }
      if (s == nsteps) break;
      // three rounds of projection against the current and previous blocks, then
      // Loewdin; repeated while directions are dropped
      int pw = p, pn = 0;
      while (true) {
        QS_TRY(project_out(wk, Vcur, p, Wb, pw, nullptr));
        QS_TRY(project_out(wk, Vcur, p, Wb, pw, nullptr));
        QS_TRY(project_out(wk, Vprev, pprev, Wb, pw, nullptr));
        QS_TRY(project_out(wk, Vprev, pprev, Wb, pw, nullptr));
        QS_TRY(project_out(wk, Vcur, p, Wb, pw, nullptr));
        QS_TRY(project_out(wk, Vprev, pprev, Wb, pw, nullptr));
        QS_TRY(lowdin(wk, Wb, pw, Vnext, sv_cut, &pn));
        if (pn == pw || pn == 0) break;
        // re-run on the reduced block: copy Vnext (pn wide) back into Wb layout
        QS_HIP(ctx, hipMemcpyAsync(Wb, Vnext, (size_t)n * pn * 16, hipMemcpyDeviceToDevice, ctx->stream));
        pw = pn;
      }
      kept[s + 1] = pn;
      if (pn == 0) break;  // invariant subspace: remaining blocks stay zero
      iprev = icur;
      icur = inext;
      inext = 3 - iprev - icur;
      Vprev = bufs[iprev];
      pprev = p;
      Vcur = bufs[icur];
      Vnext = bufs[inext];
      p = pn;
    }
    return QS_OK;
  };
  st = run();
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(mem);
  if (st != QS_OK) return st;
  memcpy(A_out, Aall.data(), Aall.size() * 16);
  memcpy(B_out, Ball.data(), Ball.size() * 16);
  if (kept_out) memcpy(kept_out, kept.data(), kept.size() * sizeof(int));
  return QS_OK;
}

extern "C" qs_status qs_block_cf_spectrum(qs_ctx *ctx, int k, int n, const double *A, const double *B,
                                          const int *kept, double E0, double emin, double emax, int ne,
                                          double gamma, double *out) {
  if (!ctx || k < 1 || k > QS_KMAX || n < 1 || ne < 1 || !A || !B || !kept || !out) return QS_EINVAL;
  if (kept[1] < 1) return qs_fail(ctx, QS_EINVAL, "empty start block");
  QS_HIP(ctx, hipSetDevice(ctx->device));
  // levels past an invariant subspace (kept == 0) are dropped
  int nl = n;
  for (int l = 1; l <= n; ++l)
    if (kept[l] == 0) {
      nl = l - 1;
      break;
    }
  std::vector<int> kp(nl + 2, 0);
  for (int l = 0; l <= nl; ++l) kp[l] = kept[l];
  const size_t kk = (size_t)k * k;
  const size_t na = (size_t)(n + 1) * kk, nb = (size_t)n * kk, no = (size_t)(ne + 1) * kk;
  char *mem = nullptr;
  QS_TRY(qs_malloc(ctx, (void **)&mem, (na + nb + no) * 16 + kp.size() * 4 + 64));
  double2 *dA = (double2 *)mem, *dB = dA + na, *dO = dB + nb;
  int *dk = (int *)(dO + no);
  hipError_t e = hipMemcpyAsync(dA, A, na * 16, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dB, B, nb * 16, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dk, kp.data(), kp.size() * 4, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_block_cf, dim3(ne + 1), dim3(QS_BLOCK), 0, ctx->stream, k, nl, dA, dB, dk, E0, emin,
                       (emax - emin) / ne, gamma, dO);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, dO, no * 16, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(mem);
  if (e != hipSuccess) return qs_hip_fail(ctx, e, "qs_block_cf_spectrum", __FILE__, __LINE__);
  return QS_OK;
}

// ---------------------------------------------------------------------------
// lowest eigenstates
// ---------------------------------------------------------------------------
extern "C" qs_status qs_ground_states(qs_ctx *ctx, const qs_hop *h, int nstates, double tol, int max_krylov,
                                      int max_restarts, uint64_t seed, const double *start, double *E_out,
                                      double *vecs, double *variance_out, int *iterations_out) {
  if (!ctx || !h || nstates < 1 || nstates > QS_KMAX || !E_out || !vecs) return QS_EINVAL;
  if (h->rows != h->cols) return qs_fail(ctx, QS_EINVAL, "qs_ground_states needs rows == cols");
  QS_HIP(ctx, hipSetDevice(ctx->device));
  const int64_t n = h->rows->n;
  if (n < nstates) return qs_fail(ctx, QS_EINVAL, "basis smaller than the number of states");
  const int p = nstates;
  if (max_krylov <= 0) max_krylov = 500;
  if (max_restarts <= 0) max_restarts = 100;
  if (tol <= 0) tol = 100.0 * std::sqrt(2.220446049250313e-16);
  const int keep_max = std::min(QS_KMAX, std::max(2 * p, p + 6));
  int mmax = (int)std::min<int64_t>(n, std::max(max_krylov, keep_max + 3 * p));
  // blocks of the Krylov basis live in one buffer, each block at a fixed slot
  const size_t slot_bytes = (size_t)n * QS_KMAX * 16;
  // storage: mmax + 2*QS_KMAX vectors of basis + W + Y + tmp
  const size_t vbytes = (size_t)n * 16;
  char *mem = nullptr;
  QS_TRY(qs_malloc(ctx, (void **)&mem, (size_t)(mmax + QS_KMAX) * vbytes + 2 * slot_bytes + 2 * vbytes + 256));
  double2 *Qbuf = (double2 *)mem;
  double2 *Wb = (double2 *)(mem + (size_t)(mmax + QS_KMAX) * vbytes);
  double2 *Yb = (double2 *)((char *)Wb + slot_bytes);
  double2 *tmp = (double2 *)((char *)Yb + slot_bytes);
  Work wk;
  qs_status st = wk.init(ctx, n);
  struct Blk {
    double2 *ptr;
    int w;
    int off;  // column offset in the projected matrix
  };
  std::vector<Blk> blocks;
  std::vector<double> theta;
  std::vector<cd> S;
  int iters = 0;
  auto run = [&]() -> qs_status {
    QS_TRY(st);
    // start block
    if (start) {
      hipLaunchKernelGGL(k_pack, dim3(grid_for(ctx, n * p)), dim3(QS_BLOCK), 0, ctx->stream, n, p,
                         (const double2 *)start, Wb, 1);
    } else {
      hipLaunchKernelGGL(k_rand, dim3(grid_for(ctx, n * p)), dim3(QS_BLOCK), 0, ctx->stream, n * p, seed, Wb);
    }
    QS_HIP(ctx, hipGetLastError());
    int w0 = 0;
    QS_TRY(lowdin(wk, Wb, p, Qbuf, 0.0, &w0));
    if (w0 == 0) return qs_fail(ctx, QS_EINVAL, "start vectors are linearly dependent");
    blocks.push_back({Qbuf, w0, 0});
    int m = w0;               // basis vectors so far
    std::vector<cd> T;        // projected matrix, column-major, ld = mmax + QS_KMAX
    const int ldT = mmax + 2 * QS_KMAX;
    T.assign((size_t)ldT * ldT, cd(0));
    auto Tat = [&](int i, int j) -> cd & { return T[(size_t)i + (size_t)j * ldT]; };
    int next_check = std::min(mmax, std::max(4 * p, 24));
    int restarts = 0;
    while (true) {
      ++iters;
      const Blk cur = blocks.back();
      QS_TRY(block_apply(ctx, h, cur.ptr, cur.w, cur.w, Wb, tmp));
      // full re-orthogonalization, two passes; coefficients accumulate the
      // projected matrix column block of `cur`
      for (int pass = 0; pass < 2; ++pass) {
        for (const Blk &b : blocks) {
          std::vector<cd> C((size_t)b.w * cur.w, cd(0));
          QS_TRY(project_out(wk, b.ptr, b.w, Wb, cur.w, &C));
          // first pass overwrites (the coupling written when the block was
          // appended is recomputed here), the second adds its correction
          for (int i = 0; i < b.w; ++i)
            for (int j = 0; j < cur.w; ++j) {
              cd &tv = Tat(b.off + i, cur.off + j);
              tv = (pass == 0 ? cd(0) : tv) + C[(size_t)i * cur.w + j];
            }
        }
      }
      for (const Blk &b : blocks)
        for (int i = 0; i < b.w; ++i)
          for (int j = 0; j < cur.w; ++j) Tat(cur.off + j, b.off + i) = std::conj(Tat(b.off + i, cur.off + j));
      // new block
      int wn = 0;
      double2 *nptr = Qbuf + (size_t)m * n;
      QS_TRY(lowdin(wk, Wb, cur.w, nptr, 0.0, &wn));
      std::vector<cd> R;  // R[i][j] = <Qnew_i|W_j>
      if (wn > 0) QS_TRY(wk.bdot(nptr, wn, wn, Wb, cur.w, cur.w, R));
      const bool exhausted = wn == 0 || m + wn > n;
      const bool full = mmax < n && m + wn + p > mmax;
      if (exhausted || full || m >= next_check) {
        // Rayleigh-Ritz on the current basis (m vectors)
        std::vector<cd> Tm((size_t)m * m);
        for (int j = 0; j < m; ++j)
          for (int i = 0; i < m; ++i) Tm[i + (size_t)j * m] = 0.5 * (Tat(i, j) + std::conj(Tat(j, i)));
        theta.assign(m, 0.0);
        S.assign((size_t)m * m, cd(0));
        qs_heev(m, (const double *)Tm.data(), theta.data(), (double *)S.data());
        // residual norms |R s_last|
        std::vector<double> var(p, 0.0);
        double sumE2 = 0.0;
        for (int i = 0; i < p; ++i) sumE2 += theta[i] * theta[i];
        bool conv = true;
        for (int i = 0; i < p; ++i) {
          double r2 = 0.0;
          for (int a = 0; a < wn; ++a) {
            cd s = 0;
            for (int j = 0; j < cur.w; ++j) s += R[(size_t)a * cur.w + j] * S[(size_t)(cur.off + j) + (size_t)i * m];
            r2 += std::norm(s);
          }
          var[i] = exhausted ? 0.0 : r2;
          if (var[i] > tol * std::max(1.0, sumE2) / p) conv = false;
        }
        if (conv || exhausted || restarts >= max_restarts) {
          // Ritz vectors of the wanted states
          std::vector<cd> Y((size_t)m * p);
          for (int b = 0; b < (int)blocks.size(); ++b) {
            const Blk &bl = blocks[b];
            std::vector<cd> Cb((size_t)bl.w * p);
            for (int i = 0; i < bl.w; ++i)
              for (int j = 0; j < p; ++j) Cb[(size_t)i * p + j] = S[(size_t)(bl.off + i) + (size_t)j * m];
            QS_TRY(wk.bupd(bl.ptr, bl.w, bl.w, Cb, Yb, p, p, b == 0 ? 0.0 : 1.0));
          }
          hipLaunchKernelGGL(k_pack, dim3(grid_for(ctx, n * p)), dim3(QS_BLOCK), 0, ctx->stream, n, p, Yb,
                             (double2 *)vecs, 0);
          QS_HIP(ctx, hipGetLastError());
          for (int i = 0; i < p; ++i) {
            E_out[i] = theta[i];
            if (variance_out) variance_out[i] = var[i];
          }
          if (!conv && !exhausted) return qs_fail(ctx, QS_ENOTCONVERGED, "not converged after %d restarts", restarts);
          return QS_OK;
        }
        if (full) {
          // thick restart: keep r Ritz vectors + the new residual block
          ++restarts;
          const int r = std::min(keep_max, m - 1);
          for (int b = 0; b < (int)blocks.size(); ++b) {
            const Blk &bl = blocks[b];
            std::vector<cd> Cb((size_t)bl.w * r);
            for (int i = 0; i < bl.w; ++i)
              for (int j = 0; j < r; ++j) Cb[(size_t)i * r + j] = S[(size_t)(bl.off + i) + (size_t)j * m];
            QS_TRY(wk.bupd(bl.ptr, bl.w, bl.w, Cb, Yb, r, r, b == 0 ? 0.0 : 1.0));
          }
          // coupling of the residual block to the kept Ritz vectors: <Qnew|H y_i> = R s_last,i
          std::vector<cd> Cnew((size_t)wn * r, cd(0));
          for (int a = 0; a < wn; ++a)
            for (int i = 0; i < r; ++i) {
              cd s = 0;
              for (int j = 0; j < cur.w; ++j) s += R[(size_t)a * cur.w + j] * S[(size_t)(cur.off + j) + (size_t)i * m];
              Cnew[(size_t)a * r + i] = s;
            }
          // move: Y -> slot 0 (r vectors), Qnew -> right after
          QS_HIP(ctx, hipMemcpyAsync(Qbuf, Yb, (size_t)n * r * 16, hipMemcpyDeviceToDevice, ctx->stream));
          QS_HIP(ctx, hipMemcpyAsync(Qbuf + (size_t)r * n, nptr, (size_t)n * wn * 16, hipMemcpyDeviceToDevice,
                                     ctx->stream));
          QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
          blocks.clear();
          blocks.push_back({Qbuf, r, 0});
          blocks.push_back({Qbuf + (size_t)r * n, wn, r});
          std::fill(T.begin(), T.end(), cd(0));
          for (int i = 0; i < r; ++i) Tat(i, i) = theta[i];
          for (int a = 0; a < wn; ++a)
            for (int i = 0; i < r; ++i) {
              Tat(r + a, i) = Cnew[(size_t)a * r + i];
              Tat(i, r + a) = std::conj(Cnew[(size_t)a * r + i]);
            }
          m = r + wn;
          next_check = std::min(mmax, m + std::max(4 * p, 24));
          continue;
        }
        next_check = std::min(mmax, (int)(next_check * 1.5) + p);
      }
      // append the new block
      for (int a = 0; a < wn; ++a)
        for (int j = 0; j < cur.w; ++j) {
          Tat(m + a, cur.off + j) = R[(size_t)a * cur.w + j];
          Tat(cur.off + j, m + a) = std::conj(R[(size_t)a * cur.w + j]);
        }
      blocks.push_back({nptr, wn, m});
      m += wn;
    }
  };
  st = run();
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(mem);
  if (iterations_out) *iterations_out = iters;
  return st;
}

extern "C" qs_status qs_basis_closure(qs_ctx *ctx, const qs_op *op, const qs_basis *seed, int nres,
                                      const uint8_t *masks, const int *lo, const int *hi, int64_t max_dim,
                                      qs_basis **out) {
  return qs_fail(ctx, QS_EUNSUPPORTED, "qs_basis_closure: not built yet");
}
