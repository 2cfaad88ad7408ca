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
#include <climits>
#include <cmath>
#include <complex>
#include <numeric>
#include <cstring>
#include <vector>

#include "qs_internal.h"
#include "qs_kernels.h"

using cd = std::complex<double>;

#define QS_KMAX 16

qs_status qs_hop_apply_vectors(qs_ctx *ctx, const qs_hop *h, int k, const double *x, int64_t ldx, int64_t ldvx,
                               double *y, int64_t ldy, int64_t ldvy);
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

// Tall-skinny products against a set of basis vectors V (vector i of length n
// at vset_ptr(V, i): the first na vectors at a + i*ld, the rest at
// b + (i - na)*ld, so two separately stored blocks act as one set) and a block
// W of w <= QS_KMAX columns (entry (r, j) at W[r*ldw + j*ldj]): C = V^+ W and
// W <- beta W + V C. Each reads V once, which is what bounds a full
// re-orthogonalization.
struct VSet {
  const double2 *a;
  const double2 *b;
  int na;
  int64_t ld;
};

__device__ __forceinline__ const double2 *vset_ptr(const VSet &s, int i) {
  return i < s.na ? s.a + (int64_t)i * s.ld : s.b + (int64_t)(i - s.na) * s.ld;
}

typedef double qs_d4 __attribute__((ext_vector_type(4)));

// part[c*m*w + i*w + j] = sum over the rows of wave chunk c (rows
// [c*rw, (c+1)*rw)) of conj(V_i[r]) W_j[r].
//
// Each wave owns one chunk and T tiles of 16 vectors; the FP64 matrix core
// (v_mfma_f64_16x16x4: A 16x4 with A[i][k] in lane i + 16k, B 4x16 with
// B[k][j] in lane j + 16k, D 16x16 with D[i][j] in lane j + 16(i%4), element
// i/4) sums over 4 rows per instruction, so the rows need no cross-lane
// reduction until the chunk is done. k = lane/16 selects row rg + 4u + k at
// sub-step u, which gives each vector 64 contiguous bytes per load.
// The complex product is split into two real ones, A = Re V then A = Im V,
// against B columns in one of two forms:
//   (Re W, Im W)  -> Re C = sum Re V Re W + Im V Im W
//   (Im W, -Re W) -> Im C = sum Re V Im W - Im V Re W
// WT = 1 (w <= 8): columns j < w hold the first form of W_j, w <= j < 2w the
// second form of W_{j-w}. WT = 2 (w <= 16): tile 0 holds the first form,
// tile 1 the second, column j for W_j.
template <int WT, int T>
__global__ void __launch_bounds__(QS_BLOCK)
k_vtw(int64_t n, int m, VSet V, const double2 *__restrict__ Wm, int64_t ldw, int64_t ldj, int w, int rw,
      double2 *__restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int64_t c = (int64_t)blockIdx.y * (QS_BLOCK / 64) + (threadIdx.x >> 6);
  const int64_t r0 = c * rw;
  if (r0 >= n) return;
  const int64_t r1 = min<int64_t>(n, r0 + rw);
  const int q = lane >> 4, col = lane & 15;
  const int t0 = blockIdx.x * T;
  const int ntiles = (m + 15) >> 4;
  bool bim[WT];
  const bool bok = WT == 1 ? col < 2 * w : col < w;
  const int bj = WT == 1 ? (col >= w ? col - w : col) : col;
  if (WT == 1) bim[0] = col >= w;
  else {
    bim[0] = false;
    bim[WT - 1] = true;
  }
  const double2 *wp = Wm + (int64_t)(bok ? bj : 0) * ldj;
  const double2 *vp[T];
  bool vok[T];
#pragma unroll
  for (int t = 0; t < T; ++t) {
    const int i = (t0 + t) * 16 + col;
    vok[t] = i < m;
    vp[t] = vset_ptr(V, vok[t] ? i : 0);
  }
  qs_d4 acc[T][WT];
#pragma unroll
  for (int t = 0; t < T; ++t)
#pragma unroll
    for (int bt = 0; bt < WT; ++bt) acc[t][bt] = qs_d4{0.0, 0.0, 0.0, 0.0};
  for (int64_t rg = r0; rg < r1; rg += 16) {
    double2 wv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t r = rg + 4 * u + q;
      wv[u] = (bok && r < r1) ? wp[r * ldw] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int t = 0; t < T; ++t) {
      if (t0 + t >= ntiles) break;
      double2 av[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int64_t r = rg + 4 * u + q;
        av[u] = (vok[t] && r < r1) ? vp[t][r] : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
#pragma unroll
        for (int bt = 0; bt < WT; ++bt) {
          const double b0 = bim[bt] ? wv[u].y : wv[u].x;
          const double b1 = bim[bt] ? -wv[u].x : wv[u].y;
          acc[t][bt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u].x, b0, acc[t][bt], 0, 0, 0);
          acc[t][bt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u].y, b1, acc[t][bt], 0, 0, 0);
        }
      }
    }
  }
  if (!bok) return;
  double *out = (double *)part;
#pragma unroll
  for (int t = 0; t < T; ++t) {
    if (t0 + t >= ntiles) break;
#pragma unroll
    for (int bt = 0; bt < WT; ++bt) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = (t0 + t) * 16 + q + 4 * v;
        if (i < m) out[2 * (((size_t)c * m + i) * w + bj) + (bim[bt] ? 1 : 0)] = acc[t][bt][v];
      }
    }
  }
}

// W[r][j] = beta W[r][j] + sum_i V_i[r] C[i][j]  (C: m x WC row-major, device)
template <int WC>
__global__ void __launch_bounds__(QS_BLOCK)
k_vupd(int64_t n, int m, VSet V, const double2 *__restrict__ C, double2 *__restrict__ Wm, int64_t ldw, int64_t ldj,
       double beta) {
  constexpr int TI = 64;
  __shared__ double2 sc[TI * WC];
  for (int64_t rb = (int64_t)blockIdx.x * blockDim.x; rb < n; rb += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = rb + threadIdx.x;
    const bool act = r < n;
    double2 acc[WC];
#pragma unroll
    for (int j = 0; j < WC; ++j) {
      acc[j] = make_double2(0.0, 0.0);
      if (act && beta != 0.0) {
        const double2 y = Wm[r * ldw + j * ldj];
        acc[j] = make_double2(beta * y.x, beta * y.y);
      }
    }
    for (int i0 = 0; i0 < m; i0 += TI) {
      const int tn = min(TI, m - i0);
      __syncthreads();
      for (int e = threadIdx.x; e < tn * WC; e += blockDim.x) sc[e] = C[(size_t)i0 * WC + e];
      __syncthreads();
      if (act) {
        for (int i = 0; i < tn; ++i) {
          const double2 v = vset_ptr(V, i0 + i)[r];
#pragma unroll
          for (int j = 0; j < WC; ++j) {
            const double2 c = sc[i * WC + j];
            acc[j].x = fma(v.x, c.x, fma(-v.y, c.y, acc[j].x));
            acc[j].y = fma(v.x, c.y, fma(v.y, c.x, acc[j].y));
          }
        }
      }
    }
    if (act) {
#pragma unroll
      for (int j = 0; j < WC; ++j) Wm[r * ldw + j * ldj] = acc[j];
    }
  }
}

// Runtime width -> template instantiation (1..QS_KMAX).
#define QS_WIDTH_SWITCH(w, CALL)                                         \
  switch (w) {                                                           \
    case 1: CALL(1); break;   case 2: CALL(2); break;                     \
    case 3: CALL(3); break;   case 4: CALL(4); break;                     \
    case 5: CALL(5); break;   case 6: CALL(6); break;                     \
    case 7: CALL(7); break;   case 8: CALL(8); break;                     \
    case 9: CALL(9); break;   case 10: CALL(10); break;                   \
    case 11: CALL(11); break; case 12: CALL(12); break;                   \
    case 13: CALL(13); break; case 14: CALL(14); break;                   \
    case 15: CALL(15); break; default: CALL(16); break;                   \
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
  qs_status init(qs_ctx *c, int64_t nn) {
    ctx = c;
    n = nn;
    return QS_OK;
  }
  // buffers of the basis products (grown on demand)
  double2 *vpart = nullptr;
  size_t vpart_bytes = 0;
  double2 *vC = nullptr;
  size_t vC_bytes = 0;
  ~Work() {
    qs_free(vpart);
    qs_free(vC);
  }
  qs_status grow(double2 **p, size_t *have, size_t need) {
    if (need <= *have) return QS_OK;
    QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    qs_free(*p);
    *p = nullptr;
    *have = 0;
    QS_TRY(qs_malloc(ctx, (void **)p, need));
    *have = need;
    return QS_OK;
  }
  // vector-major sets: one block (vectors at a + i*n) or two blocks, the
  // first holding na vectors
  VSet vs(const double2 *a) const { return VSet{a, a, INT_MAX, n}; }
  VSet vs2(const double2 *a, int na, const double2 *b) const { return VSet{a, b, na, n}; }
  // C[i*w + j] = <V_i|W_j> for the m vectors of V and the w vectors of W
  // (vector-major, W_j at W + j*n)
  qs_status vtw(const VSet &V, int m, const double2 *W, int w, std::vector<cd> &C) {
    C.assign((size_t)m * w, cd(0));
    if (m == 0 || w == 0) return QS_OK;
    if (w > QS_KMAX) return qs_fail(ctx, QS_EINVAL, "block wider than %d", QS_KMAX);
    const int wt = w <= 8 ? 1 : 2;
    const int tiles_per_wave = wt == 1 ? 8 : 4;
    const int ntiles = (m + 15) / 16;
    const int xg = (ntiles + tiles_per_wave - 1) / tiles_per_wave;
    // rows per wave: about 6144 waves in all, at least 32 rows per column so
    // the partial sums stay a small fraction of the basis
    int64_t rw = std::max<int64_t>({(n * xg + 6143) / 6144, (int64_t)32 * w, 64});
    rw = (rw + 15) / 16 * 16;
    const int64_t nch = (n + rw - 1) / rw;
    const unsigned gy = (unsigned)((nch + QS_BLOCK / 64 - 1) / (QS_BLOCK / 64));
    QS_TRY(grow(&vpart, &vpart_bytes, (size_t)nch * m * w * 16));
    QS_TRY(grow(&vC, &vC_bytes, (size_t)m * w * 16));
    if (wt == 1)
      hipLaunchKernelGGL((k_vtw<1, 8>), dim3((unsigned)xg, gy), dim3(QS_BLOCK), 0, ctx->stream, n, m, V, W,
                         (int64_t)1, n, w, (int)rw, vpart);
    else
      hipLaunchKernelGGL((k_vtw<2, 4>), dim3((unsigned)xg, gy), dim3(QS_BLOCK), 0, ctx->stream, n, m, V, W,
                         (int64_t)1, n, w, (int)rw, vpart);
    hipLaunchKernelGGL(k_bdot_reduce, dim3(m * w), dim3(QS_BLOCK), 0, ctx->stream, (int)nch, m * w,
                       (const double2 *)vpart, vC);
    QS_HIP(ctx, hipGetLastError());
    QS_TRY(qs_d2h(ctx, C.data(), vC, (size_t)m * w * 16));
    QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return QS_OK;
  }
  // W = beta W + V C (C host, m x w row-major; W vector-major)
  qs_status vupd(const VSet &V, int m, const std::vector<cd> &C, double2 *W, int w, double beta) {
    if (w == 0) return QS_OK;
    if (w > QS_KMAX) return qs_fail(ctx, QS_EINVAL, "block wider than %d", QS_KMAX);
    QS_TRY(grow(&vC, &vC_bytes, (size_t)std::max(m, 1) * w * 16));
    if (m > 0) QS_TRY(qs_h2d(ctx, vC, C.data(), (size_t)m * w * 16));
#define QS_VUPD(WC)                                                                                         \
  hipLaunchKernelGGL(k_vupd<WC>, dim3(grid_for(ctx, n)), dim3(QS_BLOCK), 0, ctx->stream, n, m, V,             \
                     (const double2 *)vC, W, (int64_t)1, n, beta)
    QS_WIDTH_SWITCH(w, QS_VUPD)
#undef QS_VUPD
    QS_HIP(ctx, hipGetLastError());
    QS_HIP(ctx, hipStreamSynchronize(ctx->stream));  // vC is reused by the next call
    return QS_OK;
  }
};

// Orthonormalize the k columns of W (in place into V): overlap S = W^+ W,
// S = U diag(l) U^+, keep directions with sqrt(l) > max(sv_cut, 3 sqrt(eps) s_max)
// (the singular-value cut of ComplexWaveFunctionListSVDOrthogonalizeSameBasisOnce,
// Core/ComplexWaveFunctionArithmetic.cpp:1356-1550). Without reduction the
// symmetric transform S^-1/2 is used, otherwise U_keep diag(l^-1/2).
// Loewdin transform from the overlap S (S[i*k+j] = <W_i|W_j>): k x kept
// row-major, V = W Tm.
std::vector<cd> lowdin_transform(const std::vector<cd> &S, int k, double sv_cut, int *kept) {
  // column-major for heev: element (i,j) at i + j*k -> transpose = conj
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
  std::vector<cd> Tm((size_t)k * std::max(kp, 1), cd(0));  // row-major k x kp
  if (kp == 0) return Tm;
  if (kp == k) {
    // S^-1/2 = U diag(l^-1/2) U^+
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) {
        cd sum = 0;
        for (int m = 0; m < k; ++m) sum += U[i + (size_t)m * k] * (1.0 / std::sqrt(l[m])) * std::conj(U[j + (size_t)m * k]);
        Tm[(size_t)i * k + j] = sum;
      }
  } else {
    for (int i = 0; i < k; ++i)
      for (int jj = 0; jj < kp; ++jj) {
        const int m = keep[jj];
        Tm[(size_t)i * kp + jj] = U[i + (size_t)m * k] / std::sqrt(l[m]);
      }
  }
  return Tm;
}

// vector-major blocks (vector v at + v*n)
qs_status lowdin_v(Work &wk, const double2 *W, int k, double2 *V, double sv_cut, int *kept) {
  std::vector<cd> S;
  QS_TRY(wk.vtw(wk.vs(W), k, W, k, S));
  const std::vector<cd> Tm = lowdin_transform(S, k, sv_cut, kept);
  if (*kept == 0) return QS_OK;
  return wk.vupd(wk.vs(W), k, Tm, V, *kept, 0.0);
}

// W -= X (X^+ W) for the kx vectors of X; the coefficients X^+ W are added to
// *accum (kx x kw row-major) when given
qs_status project_out_v(Work &wk, const VSet &X, int kx, double2 *W, int kw, std::vector<cd> *accum = nullptr) {
  if (kx == 0 || kw == 0) return QS_OK;
  std::vector<cd> C;
  QS_TRY(wk.vtw(X, kx, W, kw, C));
  if (accum)
    for (size_t i = 0; i < C.size(); ++i) (*accum)[i] += C[i];
  for (auto &c : C) c = -c;
  return wk.vupd(X, kx, C, W, kw, 1.0);
}

// y = H x for k vectors laid out one after another
qs_status block_apply_v(qs_ctx *ctx, const qs_hop *h, const double2 *x, int k, double2 *y) {
  if (k == 0) return QS_OK;
  const int64_t n = h->rows->n;
  if (h->assembled) return qs_hop_apply_vectors(ctx, h, k, (const double *)x, 1, n, (double *)y, 1, n);
  for (int v = 0; v < k; ++v) QS_TRY(qs_launch_mf_ext(ctx, h, (const double *)(x + (size_t)v * n), (double *)(y + (size_t)v * n)));
  return QS_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// block Lanczos (tensor spectra)
// ---------------------------------------------------------------------------
extern "C" qs_status qs_block_lanczos(qs_ctx *ctx, const qs_hop *h, int k, const double *V0, int nsteps,
                                      double sv_cut, double *A_out, double *B_out, int *kept_out) {
  if (!ctx) return QS_EINVAL;
  if (!h || !V0 || !A_out || !B_out || nsteps < 1) return qs_fail(ctx, QS_EINVAL, "qs_block_lanczos: bad arguments");
  if (k < 1 || k > QS_KMAX)
    return qs_fail(ctx, QS_EINVAL, "qs_block_lanczos: %d start vectors; a block holds 1 to %d", k, QS_KMAX);
  if (h->rows != h->cols) return qs_fail(ctx, QS_EINVAL, "qs_block_lanczos needs rows == cols");
  QS_HIP(ctx, hipSetDevice(ctx->device));
  const int64_t n = h->rows->n;
  const size_t bb = (size_t)n * k * 16;
  char *mem = nullptr;
  QS_TRY(qs_malloc(ctx, (void **)&mem, 5 * bb + 256));
  double2 *X0 = (double2 *)mem, *B1 = (double2 *)(mem + bb), *B2 = (double2 *)(mem + 2 * bb);
  double2 *B3 = (double2 *)(mem + 3 * bb), *Wb = (double2 *)(mem + 4 * bb);
  Work wk;
  qs_status st = wk.init(ctx, n);
  const size_t kk = (size_t)k * k;
  std::vector<cd> Aall((size_t)(nsteps + 1) * kk, cd(0)), Ball((size_t)nsteps * kk, cd(0));
  std::vector<int> kept(nsteps + 1, 0);
  kept[0] = k;
  // blocks are stored one vector after another (vector v of a block at + v*n)
  auto run = [&]() -> qs_status {
    QS_TRY(st);
    QS_HIP(ctx, hipMemcpyAsync(X0, V0, bb, hipMemcpyDeviceToDevice, ctx->stream));
    int p = 0;
    QS_TRY(lowdin_v(wk, X0, k, B1, sv_cut, &p));
    kept[1] = p;
    if (p == 0) return QS_OK;
    {  // B0[i][j] = <V1_j|X0_i>
      std::vector<cd> C;
      QS_TRY(wk.vtw(wk.vs(B1), p, X0, k, C));
      for (int i = 0; i < k; ++i)
        for (int j = 0; j < p; ++j) Ball[(size_t)i * k + j] = C[(size_t)j * k + i];
    }
    double2 *bufs[3] = {B1, B2, B3};
    int icur = 0, inext = 1, iprev = -1;
    double2 *Vprev = nullptr, *Vcur = B1, *Vnext = B2;
    int pprev = 0;
    for (int s = 1; s <= nsteps; ++s) {
      QS_TRY(block_apply_v(ctx, h, Vcur, p, Wb));
      // X = [V_{s-1}, V_s]; X^+ H V_s holds B_{s-1} (first pprev rows) and A_s
      const VSet X = pprev > 0 ? wk.vs2(Vprev, pprev, Vcur) : wk.vs(Vcur);
      const int kx = pprev + p;
      std::vector<cd> C;
      QS_TRY(wk.vtw(X, kx, Wb, p, C));
      for (int i = 0; i < p; ++i)
        for (int j = 0; j < p; ++j) Aall[(size_t)s * kk + (size_t)i * k + j] = C[(size_t)(pprev + i) * p + j];
      for (int i = 0; i < pprev; ++i)
        for (int j = 0; j < p; ++j) Ball[(size_t)(s - 1) * kk + (size_t)i * k + j] = C[(size_t)i * p + j];
      if (s == nsteps) break;
      // two rounds of classical Gram-Schmidt against the previous and current
      // blocks together (the first with the coefficients above), then Loewdin;
      // repeated on the reduced block while directions are dropped
      for (auto &c : C) c = -c;
      QS_TRY(wk.vupd(X, kx, C, Wb, p, 1.0));
      int pw = p, pn = 0;
      while (true) {
        QS_TRY(project_out_v(wk, X, kx, Wb, pw));
        QS_TRY(lowdin_v(wk, Wb, pw, Vnext, sv_cut, &pn));
        if (pn == pw || pn == 0) break;
        QS_HIP(ctx, hipMemcpyAsync(Wb, Vnext, (size_t)n * pn * 16, hipMemcpyDeviceToDevice, ctx->stream));
        pw = pn;
        QS_TRY(project_out_v(wk, X, kx, Wb, pw));
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
  qs_free(mem);
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
  qs_free(mem);
  if (e != hipSuccess) return qs_hip_fail(ctx, e, "qs_block_cf_spectrum", __FILE__, __LINE__);
  return QS_OK;
}

// ---------------------------------------------------------------------------
// Lanczos with the Krylov basis kept (KrylovBasis = true)
// ---------------------------------------------------------------------------
// LanczosTriDiagonalizeKrylovCC (Core/Lanczos.cpp:3705-3854): every new vector
// H v_{i-1} is orthogonalized against all kept vectors twice (a[i] is its
// overlap with v_{i-1} before any subtraction), normalized (b[i] its norm;
// below 100 DBL_MIN it becomes zero), and zeroed when its squared overlaps
// with the kept vectors still exceed eps (the Krylov space is exhausted);
// a[nsteps] = <v_{nsteps-1}|H|v_{nsteps-1}>. Gram-Schmidt runs in classical
// form on the FP64 matrix cores (k_vtw / k_vupd), where the reference subtracts
// vector by vector; both leave the basis orthonormal to working precision.
extern "C" qs_status qs_lanczos_krylov(qs_ctx *ctx, const qs_hop *h, const double *v0, int nsteps, double eps,
                                       double *a_out, double *b_out, double *V_out) {
  if (!ctx) return QS_EINVAL;
  if (!h || !v0 || !a_out || !b_out || !V_out || nsteps < 1)
    return qs_fail(ctx, QS_EINVAL, "qs_lanczos_krylov: bad arguments");
  if (h->rows != h->cols) return qs_fail(ctx, QS_EINVAL, "qs_lanczos_krylov needs rows == cols");
  QS_HIP(ctx, hipSetDevice(ctx->device));
  const int64_t n = h->rows->n;
  double2 *V = (double2 *)V_out;
  Work wk;
  QS_TRY(wk.init(ctx, n));
  double2 *Wb = nullptr;
  QS_TRY(qs_malloc(ctx, (void **)&Wb, (size_t)std::max<int64_t>(n, 1) * 16));
  std::vector<double> a(nsteps + 1, 0.0), b(nsteps, 0.0);
  const double tiny = 100 * 2.2250738585072014e-308;
  auto vec = [&](int i) { return V + (size_t)i * n; };
  auto run = [&]() -> qs_status {
    QS_HIP(ctx, hipMemcpyAsync(vec(0), v0, (size_t)n * 16, hipMemcpyDeviceToDevice, ctx->stream));
    std::vector<cd> C;
    QS_TRY(wk.vtw(wk.vs(vec(0)), 1, vec(0), 1, C));
    b[0] = std::sqrt(std::max(C[0].real(), 0.0));
    QS_TRY(wk.vupd(wk.vs(vec(0)), 0, C, vec(0), 1, b[0] < tiny ? 0.0 : 1.0 / b[0]));
    for (int i = 1; i < nsteps; ++i) {
      double2 *w = vec(i);
      QS_TRY(block_apply_v(ctx, h, vec(i - 1), 1, w));
      const VSet K = wk.vs(vec(0));  // the kept vectors 0 .. i-1
      for (int pass = 0; pass < 2; ++pass) {
        QS_TRY(wk.vtw(K, i, w, 1, C));
        if (pass == 0) a[i] = C[(size_t)i - 1].real();
        for (auto &c : C) c = -c;
        QS_TRY(wk.vupd(K, i, C, w, 1, 1.0));
      }
      QS_TRY(wk.vtw(wk.vs(w), 1, w, 1, C));
      b[i] = std::sqrt(std::max(C[0].real(), 0.0));
      QS_TRY(wk.vupd(K, 0, C, w, 1, b[i] < tiny ? 0.0 : 1.0 / b[i]));
      if (b[i] >= tiny) {  // a vector that still overlaps the kept ones closes the space
        QS_TRY(wk.vtw(K, i, w, 1, C));
        double ov = 0.0;
        for (int j = i - 1; j >= 0; --j) {
          ov += std::norm(C[(size_t)j]);
          if (ov > eps) {
            QS_TRY(wk.vupd(K, 0, C, w, 1, 0.0));
            break;
          }
        }
      }
    }
    // a[nsteps] = <v|H|v> for the last kept vector
    QS_TRY(block_apply_v(ctx, h, vec(nsteps - 1), 1, Wb));
    QS_TRY(wk.vtw(wk.vs(vec(nsteps - 1)), 1, Wb, 1, C));
    a[nsteps] = C[0].real();
    return QS_OK;
  };
  const qs_status st = run();
  (void)hipStreamSynchronize(ctx->stream);
  qs_free(Wb);
  if (st != QS_OK) return st;
  memcpy(a_out, a.data(), a.size() * sizeof(double));
  memcpy(b_out, b.data(), b.size() * sizeof(double));
  return QS_OK;
}

// ---------------------------------------------------------------------------
// lowest eigenstates
// ---------------------------------------------------------------------------
extern "C" qs_status qs_ground_states_ex(qs_ctx *ctx, const qs_hop *h, int nstates, double tol, int krylov_start,
                                         int krylov_step, int max_krylov, int max_restarts, uint64_t seed,
                                         int flags, const double *start, double *E_out, double *vecs,
                                         double *variance_out, int *iterations_out) {
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
  if (krylov_start <= 0) krylov_start = 100;
  if (krylov_step < 0) krylov_step = 10;
  // convergence tested only when the basis reaches its restart size
  const bool cycle_end = (flags & QS_GS_CYCLE_END) != 0;
  const int mmax = (int)std::min<int64_t>(n, std::max(max_krylov, keep_max + 3 * p));
  // basis size before a thick restart: nstates + krylov_start at first, grown
  // by krylov_step + nstates per restart while below max_krylov
  int cap = std::min(mmax, std::max(p + krylov_start, keep_max + 3 * p));
  // the Krylov basis and the working blocks are stored one vector after
  // another (vector i at Q + i*n), so one pass over the basis projects a new
  // block against every basis vector
  const size_t vbytes = (size_t)n * 16;
  const size_t slot_bytes = (size_t)n * QS_KMAX * 16;
  char *mem = nullptr;
  QS_TRY(qs_malloc(ctx, (void **)&mem, (size_t)(mmax + QS_KMAX) * vbytes + 3 * slot_bytes + 256));
  double2 *Q = (double2 *)mem;
  double2 *Wb = (double2 *)(mem + (size_t)(mmax + QS_KMAX) * vbytes);
  double2 *Yb = (double2 *)((char *)Wb + slot_bytes);
  double2 *Cb = (double2 *)((char *)Yb + slot_bytes);
  Work wk;
  qs_status st = wk.init(ctx, n);
  struct Blk {
    int off;  // first basis vector (and column of the projected matrix)
    int w;
  };
  std::vector<Blk> blocks;
  std::vector<double> theta;
  std::vector<cd> S;
  // block of w vectors -> basis vectors [off, off + w)
  auto store = [&](const double2 *blk, int w, int off) -> qs_status {
    QS_HIP(ctx, hipMemcpyAsync(Q + (size_t)off * n, blk, (size_t)w * n * 16, hipMemcpyDeviceToDevice, ctx->stream));
    return QS_OK;
  };
  int iters = 0;
  auto run = [&]() -> qs_status {
    QS_TRY(st);
    // start block
    if (start) {
      QS_HIP(ctx, hipMemcpyAsync(Wb, start, (size_t)n * p * 16, hipMemcpyDeviceToDevice, ctx->stream));
    } else {
      hipLaunchKernelGGL(k_rand, dim3(grid_for(ctx, n * p)), dim3(QS_BLOCK), 0, ctx->stream, n * p, seed, Wb);
    }
    QS_HIP(ctx, hipGetLastError());
    int w0 = 0;
    QS_TRY(lowdin_v(wk, Wb, p, Cb, 0.0, &w0));
    if (w0 == 0) return qs_fail(ctx, QS_EINVAL, "start vectors are linearly dependent");
    QS_TRY(store(Cb, w0, 0));
    blocks.push_back({0, w0});
    int m = w0;               // basis vectors so far
    std::vector<cd> T;        // projected matrix, column-major
    const int ldT = mmax + 2 * QS_KMAX;
    T.assign((size_t)ldT * ldT, cd(0));
    auto Tat = [&](int i, int j) -> cd & { return T[(size_t)i + (size_t)j * ldT]; };
    int next_check = std::min(cap, std::max(4 * p, 24));
    int restarts = 0;
    while (true) {
      ++iters;
      const Blk cur = blocks.back();
      QS_TRY(block_apply_v(ctx, h, Q + (size_t)cur.off * n, cur.w, Wb));
      // Local step: project W = H Q_cur against the current block and the one
      // before it (basis range [lo, m)), which hold all of its components in
      // exact arithmetic; the same product gives W^+ W for the norms below.
      const int lo = blocks.size() >= 2 ? blocks[blocks.size() - 2].off : cur.off;
      const int nl = m - lo;
      std::vector<cd> C;
      QS_TRY(wk.vtw(wk.vs2(Q + (size_t)lo * n, nl, Wb), nl + cur.w, Wb, cur.w, C));
      std::vector<double> wnorm2(cur.w);
      for (int j = 0; j < cur.w; ++j) {
        double r2 = C[(size_t)(nl + j) * cur.w + j].real();
        for (int i = 0; i < nl; ++i) r2 -= std::norm(C[(size_t)i * cur.w + j]);
        wnorm2[j] = r2;
      }
      C.resize((size_t)nl * cur.w);
      for (int i = 0; i < nl; ++i)
        for (int j = 0; j < cur.w; ++j) Tat(lo + i, cur.off + j) = C[(size_t)i * cur.w + j];
      for (auto &c : C) c = -c;
      QS_TRY(wk.vupd(wk.vs(Q + (size_t)lo * n), nl, C, Wb, cur.w, 1.0));
      // Full re-orthogonalization against all m basis vectors: one pass, and a
      // second when the first removed more than half of a column's norm
      // (Daniel-Gragg-Kaufman-Stewart). Its coefficients on the local range
      // correct the projected matrix; those on older vectors are rounding-level
      // and are dropped, which keeps the projected matrix block tridiagonal
      // (plus the restart's arrow).
      for (int pass = 0; pass < 2; ++pass) {
        QS_TRY(wk.vtw(wk.vs(Q), m, Wb, cur.w, C));
        bool again = false;
        for (int j = 0; j < cur.w; ++j) {
          double rem = 0.0;
          for (int i = 0; i < m; ++i) rem += std::norm(C[(size_t)i * cur.w + j]);
          if (rem > 0.5 * wnorm2[j]) again = true;
          wnorm2[j] -= rem;
        }
        for (int i = lo; i < m; ++i)
          for (int j = 0; j < cur.w; ++j) Tat(i, cur.off + j) += C[(size_t)i * cur.w + j];
        for (auto &c : C) c = -c;
        QS_TRY(wk.vupd(wk.vs(Q), m, C, Wb, cur.w, 1.0));
        if (!again) break;
      }
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < cur.w; ++j) Tat(cur.off + j, i) = std::conj(Tat(i, cur.off + j));
      // new block (in Cb; the current block is already stored)
      int wn = 0;
      QS_TRY(lowdin_v(wk, Wb, cur.w, Cb, 0.0, &wn));
      std::vector<cd> R;  // R[i][j] = <Qnew_i|W_j>
      if (wn > 0) QS_TRY(wk.vtw(wk.vs(Cb), wn, Wb, cur.w, R));
      const bool exhausted = wn == 0 || m + wn > n;
      const bool full = cap < n && m + wn + p > cap;
      if (exhausted || full || (!cycle_end && m >= next_check)) {
        // Rayleigh-Ritz on the current basis (m vectors)
        std::vector<cd> Tm((size_t)m * m);
        for (int j = 0; j < m; ++j)
          for (int i = 0; i < m; ++i) Tm[i + (size_t)j * m] = 0.5 * (Tat(i, j) + std::conj(Tat(j, i)));
        // all Ritz values; vectors of the lowest max(p, keep_max) only
        const int nvec = std::min(m, std::max(p, keep_max));
        theta.assign(m, 0.0);
        S.assign((size_t)m * nvec, cd(0));
        if (!qs_heev_lowest(m, (const double *)Tm.data(), theta.data(), (double *)S.data(), nvec))
          return qs_fail(ctx, QS_ENOTCONVERGED, "projected eigenproblem (%d x %d) did not converge", m, m);
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
        // the first `cnt` Ritz vectors into Yb: Y = Q S[:, :cnt]
        auto ritz = [&](int cnt) -> qs_status {
          std::vector<cd> Cm((size_t)m * cnt);
          for (int i = 0; i < m; ++i)
            for (int j = 0; j < cnt; ++j) Cm[(size_t)i * cnt + j] = S[(size_t)i + (size_t)j * m];
          return wk.vupd(wk.vs(Q), m, Cm, Yb, cnt, 0.0);
        };
        if (conv || exhausted || restarts >= max_restarts) {
          QS_TRY(ritz(p));
          QS_HIP(ctx, hipMemcpyAsync(vecs, Yb, (size_t)n * p * 16, hipMemcpyDeviceToDevice, ctx->stream));
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
          QS_TRY(ritz(r));
          // coupling of the residual block to the kept Ritz vectors: <Qnew|H y_i> = R s_last,i
          std::vector<cd> Cnew((size_t)wn * r, cd(0));
          for (int a = 0; a < wn; ++a)
            for (int i = 0; i < r; ++i) {
              cd s = 0;
              for (int j = 0; j < cur.w; ++j) s += R[(size_t)a * cur.w + j] * S[(size_t)(cur.off + j) + (size_t)i * m];
              Cnew[(size_t)a * r + i] = s;
            }
          // basis: the r Ritz vectors, then the residual block
          QS_TRY(store(Yb, r, 0));
          QS_TRY(store(Cb, wn, r));
          blocks.clear();
          blocks.push_back({0, r});
          blocks.push_back({r, wn});
          std::fill(T.begin(), T.end(), cd(0));
          for (int i = 0; i < r; ++i) Tat(i, i) = theta[i];
          for (int a = 0; a < wn; ++a)
            for (int i = 0; i < r; ++i) {
              Tat(r + a, i) = Cnew[(size_t)a * r + i];
              Tat(i, r + a) = std::conj(Cnew[(size_t)a * r + i]);
            }
          m = r + wn;
          if (cap + krylov_step + p < max_krylov) cap = std::min(mmax, cap + krylov_step + p);
          next_check = std::min(cap, m + std::max(4 * p, 24));
          continue;
        }
        next_check = std::min(cap, (int)(next_check * 1.5) + p);
      }
      // append the new block
      for (int a = 0; a < wn; ++a)
        for (int j = 0; j < cur.w; ++j) {
          Tat(m + a, cur.off + j) = R[(size_t)a * cur.w + j];
          Tat(cur.off + j, m + a) = std::conj(R[(size_t)a * cur.w + j]);
        }
      QS_TRY(store(Cb, wn, m));
      blocks.push_back({m, wn});
      m += wn;
    }
  };
  st = run();
  (void)hipStreamSynchronize(ctx->stream);
  qs_free(mem);
  if (iterations_out) *iterations_out = iters;
  return st;
}

extern "C" qs_status qs_ground_states(qs_ctx *ctx, const qs_hop *h, int nstates, double tol, int max_krylov,
                                      int max_restarts, uint64_t seed, const double *start, double *E_out,
                                      double *vecs, double *variance_out, int *iterations_out) {
  return qs_ground_states_ex(ctx, h, nstates, tol, 100, 10, max_krylov, max_restarts, seed, 0, start, E_out, vecs,
                             variance_out, iterations_out);
}


