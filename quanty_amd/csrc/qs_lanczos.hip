// Vector reductions, the scalar Lanczos recurrence on the device, and the
// continued-fraction Green's function.
//
// Lanczos step i (v_i = s_i u_i is kept as an unnormalized vector u_i and a
// scalar s_i = 1/b_i, so no separate scaling pass is needed):
//   K1: h = H u_i;  a_{i+1} = s_i^2 Re<h|u_i>;  w = s_i h - b_i s_{i-1} u_{i-1}
//       (w overwrites u_{i-1} in place; the Re<h|u_i> partials stay per block)
//   K2: u_{i+1} = w - a_{i+1} s_i u_i (in place);  partials of |u_{i+1}|^2
// Each kernel's prologue sums the other kernel's per-block partials in a fixed
// order, so every block sees the same scalar and results are deterministic.
// This is the recurrence of LanczosTriDiagonalizeCC (Core/Lanczos.cpp:2169-2264)
// with epsilon = 0, on a fixed basis.
#include <algorithm>
#include <cmath>
#include <vector>

#include "qs_internal.h"
#include "qs_kernels.h"

#define QS_RED_BLOCKS 1024

qs_status qs_launch_mf_ext(qs_ctx *ctx, const qs_hop *h, const double *x, double *y);

// ---------------------------------------------------------------------------
// dot / norm partials
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(QS_BLOCK)
k_dot_partial(int64_t n, const double2 *__restrict__ a, const double2 *__restrict__ b,
              double *__restrict__ pre, double *__restrict__ pim) {
  __shared__ double red[QS_BLOCK];
  double sr = 0.0, si = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 x = a[i], y = b[i];
    sr += x.x * y.x + x.y * y.y;
    si += x.x * y.y - x.y * y.x;
  }
  const double r = block_sum_values(sr, red);
  const double m = block_sum_values(si, red);
  if (threadIdx.x == 0) {
    pre[blockIdx.x] = r;
    pim[blockIdx.x] = m;
  }
}

static int red_blocks(qs_ctx *ctx, int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(QS_RED_BLOCKS, (n + QS_BLOCK - 1) / QS_BLOCK));
}

qs_status qs_dot_internal(qs_ctx *ctx, int64_t n, const double *a, const double *b, double *re, double *im) {
  const int nb = red_blocks(ctx, n);
  void *scr;
  QS_TRY(qs_scratch(ctx, (size_t)nb * 16, &scr));
  double *p = (double *)scr;
  hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(QS_BLOCK), 0, ctx->stream, n, (const double2 *)a,
                     (const double2 *)b, p, p + nb);
  QS_HIP(ctx, hipGetLastError());
  std::vector<double> h((size_t)nb * 2);
  QS_TRY(qs_d2h(ctx, h.data(), p, (size_t)nb * 16));
  QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  double sr = 0.0, si = 0.0;
  for (int i = 0; i < nb; ++i) {
    sr += h[i];
    si += h[nb + i];
  }
  *re = sr;
  *im = si;
  return QS_OK;
}

extern "C" qs_status qs_dot(qs_ctx *ctx, int64_t n, const double *a, const double *b, double *re, double *im) {
  if (!ctx || !re || !im || n < 0) return QS_EINVAL;
  QS_HIP(ctx, hipSetDevice(ctx->device));
  return qs_dot_internal(ctx, n, a, b, re, im);
}

extern "C" qs_status qs_norm(qs_ctx *ctx, int64_t n, const double *a, double *nrm) {
  if (!ctx || !nrm || n < 0) return QS_EINVAL;
  double re, im;
  QS_HIP(ctx, hipSetDevice(ctx->device));
  QS_TRY(qs_dot_internal(ctx, n, a, a, &re, &im));
  *nrm = std::sqrt(re);
  return QS_OK;
}

// ---------------------------------------------------------------------------
// scalar Lanczos
// ---------------------------------------------------------------------------
struct LzState {
  double *a;   // n+1
  double *b;   // n
  double *s;   // n (1/b)
  double *pn;  // norm partials (nb2)
  double *pd;  // dot partials (nb1)
};

__global__ void __launch_bounds__(QS_BLOCK)
k_lz_init(int64_t n, const double2 *__restrict__ v0, double2 *__restrict__ uc, double *__restrict__ pn) {
  __shared__ double red[QS_BLOCK];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 x = v0[i];
    uc[i] = x;
    s += x.x * x.x + x.y * x.y;
  }
  const double r = block_sum_values(s, red);
  if (threadIdx.x == 0) pn[blockIdx.x] = r;
}

// prologue shared by both K1 variants: b_i = |u_i|, s_i = 1/b_i
__device__ __forceinline__ void lz_k1_prologue(int step, int nb2, LzState S, double *red, double &sc, double &coef_p) {
  const double nrm2 = block_sum_fixed(S.pn, nb2, red);
  const double bc = sqrt(nrm2);
  sc = bc > 0.0 ? 1.0 / bc : 0.0;
  const double sp = step > 0 ? S.s[step - 1] : 0.0;
  coef_p = step > 0 ? bc * sp : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S.b[step] = bc;
    S.s[step] = sc;
  }
}

// K1, assembled operator (fused SpMV + epilogue)
template <int NW, bool PACKED, bool LDS>
// the packed one-word variant fits 64 VGPRs (8 waves per SIMD) without spills
__global__ void __launch_bounds__(QS_BLOCK, (NW == 1 && PACKED) ? 8 : QS_MINB_WIDE)
k_lz1_asm(int step, int nb2, LzState S, SellD A, const double2 *__restrict__ uc, double2 *__restrict__ up) {
  __shared__ double red[QS_BLOCK / 64];
  extern __shared__ char smem[];
  const double2 *tv = A.val;
  const uint64_t *tp = A.par;
  if (LDS) {
    stage_term_tables<NW>(A, smem);
    tv = reinterpret_cast<const double2 *>(smem);
    tp = reinterpret_cast<const uint64_t *>(smem + (size_t)A.nterms * 16);
  }
  double sc, cp;
  lz_k1_prologue(step, nb2, S, red, sc, cp);  // ends with a barrier
  const int lane = threadIdx.x & 63;
  int64_t s0, s1, ds;
  xcd_items(A.nslices, s0, s1, ds);
  s0 = wave_uniform(s0);
  double dot = 0.0;
  for (int64_t s = s0; s < s1; s += ds) {
    double hr[1], hi[1];
    sell_row<1, NW, PACKED>(A, s, lane, tv, tp, uc, 1, 0, hr, hi);
    const int64_t r = s * 64 + lane;
    if (r < A.nrows) {
      const double2 u = uc[r];
      dot += hr[0] * u.x + hi[0] * u.y;
      const double2 p = up[r];
      up[r] = make_double2(sc * hr[0] - cp * p.x, sc * hi[0] - cp * p.y);
    }
  }
  const double t = block_sum_values(dot, red);
  if (threadIdx.x == 0) S.pd[blockIdx.x] = t;
}

// K1, product layout (fused apply in the binding's row order + epilogue)
template <bool LDS>
__global__ void __launch_bounds__(QS_BLOCK)
k_lz1_prod(int step, int nb2, LzState S, SellD A, const double2 *__restrict__ uc, double2 *__restrict__ up) {
  __shared__ double red[QS_BLOCK / 64];
  extern __shared__ char smem[];
  const double2 *tv = A.val;
  if (LDS) {
    stage_term_values(A, smem);
    tv = reinterpret_cast<const double2 *>(smem);
  }
  double2 *xs = prod_stage_area(smem, A.nterms, LDS, 1);
  double sc, cp;
  lz_k1_prologue(step, nb2, S, red, sc, cp);  // ends with a barrier
  const int lane = threadIdx.x & 63;
  int64_t s0, s1, ds;
  xcd_items(A.nslices, s0, s1, ds);
  s0 = wave_uniform(s0);
  double dot = 0.0;
  for (int64_t s = s0; s < s1; s += ds) {
    double hr[1], hi[1];
    const int64_t r = prod_row<1>(A, s, lane, tv, uc, 1, 0, hr, hi, xs);
    if (r >= 0) {
      const double2 u = uc[r];
      dot += hr[0] * u.x + hi[0] * u.y;
      const double2 p = up[r];
      up[r] = make_double2(sc * hr[0] - cp * p.x, sc * hi[0] - cp * p.y);
    }
  }
  const double t = block_sum_values(dot, red);
  if (threadIdx.x == 0) S.pd[blockIdx.x] = t;
}

// Kernel, dynamic LDS and grid for K1 on an assembled operator.
struct Lz1Launch {
  const void *kern;
  size_t shm;
  int blocks;
};

template <int NW, bool PACKED>
static Lz1Launch lz1_choose(qs_ctx *ctx, const SellD &A) {
  Lz1Launch L;
  L.shm = term_table_bytes(A.nterms, NW);
  if (L.shm <= (size_t)ctx->term_lds) {
    L.kern = (const void *)k_lz1_asm<NW, PACKED, true>;
  } else {
    L.kern = (const void *)k_lz1_asm<NW, PACKED, false>;
    L.shm = 0;
  }
  // grid stride over slices with no more workgroups than stay resident
  L.blocks = (int)std::max<int64_t>(1, std::min<int64_t>((A.nslices + 3) / 4, qs_resident_blocks(ctx, L.kern, L.shm)));
  return L;
}

static Lz1Launch lz1_launch_for(qs_ctx *ctx, const SellD &A) {
  if (A.kind == 1) {
    Lz1Launch L;
    if (term_value_bytes(A.nterms) <= (size_t)ctx->term_lds) {
      L.kern = (const void *)k_lz1_prod<true>;
      L.shm = prod_lds_bytes(A.nterms, true, 1);
    } else {
      L.kern = (const void *)k_lz1_prod<false>;
      L.shm = prod_lds_bytes(A.nterms, false, 1);
    }
    L.blocks = (int)std::max<int64_t>(1, std::min<int64_t>((A.nslices + 3) / 4, qs_resident_blocks(ctx, L.kern, L.shm)));
    return L;
  }
  if (A.nw == 1) return A.packed ? lz1_choose<1, true>(ctx, A) : lz1_choose<1, false>(ctx, A);
  return A.packed ? lz1_choose<2, true>(ctx, A) : lz1_choose<2, false>(ctx, A);
}

// K1 epilogue for the matrix-free operator (h already in hbuf)
__global__ void __launch_bounds__(QS_BLOCK)
k_lz1_epi(int step, int nb2, LzState S, int64_t n, const double2 *__restrict__ hbuf,
          const double2 *__restrict__ uc, double2 *__restrict__ up) {
  __shared__ double red[QS_BLOCK];
  double sc, cp;
  lz_k1_prologue(step, nb2, S, red, sc, cp);
  double dot = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    const double2 h = hbuf[r];
    const double2 u = uc[r];
    dot += h.x * u.x + h.y * u.y;
    const double2 p = up[r];
    up[r] = make_double2(sc * h.x - cp * p.x, sc * h.y - cp * p.y);
  }
  const double t = block_sum_values(dot, red);
  if (threadIdx.x == 0) S.pd[blockIdx.x] = t;
}

// K2: a_{i+1}, u_{i+1} = w - a s_i u_i, partial norms
__global__ void __launch_bounds__(QS_BLOCK)
k_lz2(int step, int nb1, LzState S, int64_t n, const double2 *__restrict__ uc, double2 *__restrict__ up) {
  __shared__ double red[QS_BLOCK];
  const double sc = S.s[step];
  const double a = sc * sc * block_sum_fixed(S.pd, nb1, red);
  if (blockIdx.x == 0 && threadIdx.x == 0) S.a[step + 1] = a;
  const double f = a * sc;
  double nrm = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    const double2 w = up[r];
    const double2 u = uc[r];
    const double2 un = make_double2(w.x - f * u.x, w.y - f * u.y);
    up[r] = un;
    nrm += un.x * un.x + un.y * un.y;
  }
  const double t = block_sum_values(nrm, red);
  if (threadIdx.x == 0) S.pn[blockIdx.x] = t;
}

extern "C" qs_status qs_lanczos(qs_ctx *ctx, const qs_hop *h, const double *v0, int nsteps, double *a_out,
                                double *b_out) {
  if (!ctx || !h || !a_out || !b_out || nsteps < 1) return QS_EINVAL;
  if (h->rows != h->cols) return qs_fail(ctx, QS_EINVAL, "qs_lanczos needs rows == cols");
  QS_HIP(ctx, hipSetDevice(ctx->device));
  const int64_t n = h->rows->n;
  if (n == 0) return qs_fail(ctx, QS_EINVAL, "empty basis");
  const int nb2 = red_blocks(ctx, n);
  int nb1 = nb2;
  Lz1Launch l1{nullptr, 0, 0};
  if (h->assembled) {
    l1 = lz1_launch_for(ctx, h->view);
    nb1 = l1.blocks;
  }
  const size_t vbytes = (size_t)n * 16;
  const size_t sbytes = ((size_t)(nsteps + 1) + 2 * (size_t)nsteps + nb1 + nb2 + 64) * 8;
  char *mem = nullptr;
  const size_t need = 2 * vbytes + (h->assembled ? 0 : vbytes) + sbytes + 1024;
  QS_TRY(qs_malloc(ctx, (void **)&mem, need));
  double2 *u0 = (double2 *)mem;
  double2 *u1 = (double2 *)(mem + vbytes);
  double2 *hb = h->assembled ? nullptr : (double2 *)(mem + 2 * vbytes);
  double *sc = (double *)(mem + (h->assembled ? 2 : 3) * vbytes);
  LzState S;
  S.a = sc;
  S.b = S.a + nsteps + 1;
  S.s = S.b + nsteps;
  S.pn = S.s + nsteps;
  S.pd = S.pn + nb2;
  qs_status st = QS_OK;
  hipError_t e = hipMemsetAsync(S.a, 0, (size_t)(nsteps + 1) * 8, ctx->stream);
  // a product-layout binding runs in its own row order: v0 goes there first
  const double2 *vs = (const double2 *)v0;
  if (e == hipSuccess && h->view.kind == 1) {
    st = qs_prod_to_rows(ctx, h, 1, (const double2 *)v0, n, u1, n);
    vs = u1;
  }
  if (e == hipSuccess && st == QS_OK) {
    hipLaunchKernelGGL(k_lz_init, dim3(nb2), dim3(QS_BLOCK), 0, ctx->stream, n, vs, u0, S.pn);
    e = hipGetLastError();
  }
  if (e == hipSuccess && st == QS_OK) e = hipMemsetAsync(u1, 0, vbytes, ctx->stream);
  double2 *uc = u0, *up = u1;
  for (int i = 0; i < nsteps && e == hipSuccess && st == QS_OK; ++i) {
    if (h->assembled) {
      void *args[] = {&i, (void *)&nb2, &S, (void *)&h->view, &uc, &up};
      e = hipLaunchKernel(l1.kern, dim3(nb1), dim3(QS_BLOCK), args, l1.shm, ctx->stream);
      if (e != hipSuccess) break;
    } else {
      // matrix-free: h = H u_c into hb, then the epilogue
      st = qs_launch_mf_ext(ctx, h, (const double *)uc, (double *)hb);
      if (st != QS_OK) break;
      hipLaunchKernelGGL(k_lz1_epi, dim3(nb1), dim3(QS_BLOCK), 0, ctx->stream, i, nb2, S, n,
                         (const double2 *)hb, (const double2 *)uc, up);
    }
    e = hipGetLastError();
    if (e != hipSuccess) break;
    hipLaunchKernelGGL(k_lz2, dim3(nb2), dim3(QS_BLOCK), 0, ctx->stream, i, nb1, S, n, (const double2 *)uc, up);
    e = hipGetLastError();
    std::swap(uc, up);
  }
  if (e == hipSuccess && st == QS_OK) st = qs_d2h(ctx, a_out, S.a, (size_t)(nsteps + 1) * 8);
  if (e == hipSuccess && st == QS_OK) st = qs_d2h(ctx, b_out, S.b, (size_t)nsteps * 8);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  qs_free(mem);
  if (st != QS_OK) return st;
  if (e != hipSuccess) return qs_hip_fail(ctx, e, "qs_lanczos", __FILE__, __LINE__);
  a_out[0] = 0.0;
  return QS_OK;
}

// ---------------------------------------------------------------------------
// continued fraction (TriDiagonalMatrixToSpectrum, Core/Spectral.cpp:1433-1480)
// ---------------------------------------------------------------------------
__global__ void k_cf(int nspec, int n, const double *__restrict__ a, const double *__restrict__ b,
                     const double *__restrict__ E0, double emin, double de, int ne, double gamma,
                     double *__restrict__ re, double *__restrict__ im) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t np = (int64_t)(ne + 1);
  if (idx >= np * nspec) return;
  const int sp = (int)(idx / np);
  const int i = (int)(idx % np);
  const double *as = a + (size_t)sp * (n + 1);
  const double *bs = b + (size_t)sp * n;
  const double hnu = (double)i * de + emin + E0[sp];
  const double g2 = gamma * 0.5;
  double den = g2 * g2 + (hnu - as[n]) * (hnu - as[n]);
  double vr = (hnu - as[n]) / den;
  double vi = -g2 / den;
  for (int j = (n > 1 ? n : 1) - 1; j >= 1; --j) {
    const double b2 = bs[j] * bs[j];
    const double xr = hnu - as[j] - vr * b2;
    const double xi = g2 - vi * b2;
    den = xi * xi + xr * xr;
    vr = xr / den;
    vi = (vi * b2 - g2) / den;
  }
  re[idx] = as[0] + vr * bs[0] * bs[0];
  im[idx] = vi * bs[0] * bs[0];
}

extern "C" qs_status qs_cf_spectrum(qs_ctx *ctx, int nspec, int n, const double *a, const double *b,
                                    const double *E0, double emin, double emax, int ne, double gamma,
                                    double *re, double *im) {
  if (!ctx || nspec < 1 || n < 1 || ne < 1 || !a || !b || !E0 || !re || !im) return QS_EINVAL;
  QS_HIP(ctx, hipSetDevice(ctx->device));
  const size_t na = (size_t)nspec * (n + 1), nbv = (size_t)nspec * n, np = (size_t)nspec * (ne + 1);
  char *mem = nullptr;
  QS_TRY(qs_malloc(ctx, (void **)&mem, (na + nbv + nspec + 2 * np) * 8));
  double *da = (double *)mem, *db = da + na, *de0 = db + nbv, *dre = de0 + nspec, *dim = dre + np;
  qs_status st = qs_h2d(ctx, da, a, na * 8);
  if (st == QS_OK) st = qs_h2d(ctx, db, b, nbv * 8);
  if (st == QS_OK) st = qs_h2d(ctx, de0, E0, (size_t)nspec * 8);
  hipError_t e = hipSuccess;
  if (st == QS_OK) {
    const double de = (emax - emin) / (double)ne;
    hipLaunchKernelGGL(k_cf, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, ctx->stream, nspec, n, da, db, de0,
                       emin, de, ne, gamma, dre, dim);
    e = hipGetLastError();
  }
  if (st == QS_OK && e == hipSuccess) st = qs_d2h(ctx, re, dre, np * 8);
  if (st == QS_OK && e == hipSuccess) st = qs_d2h(ctx, im, dim, np * 8);
  (void)hipStreamSynchronize(ctx->stream);
  qs_free(mem);
  if (st != QS_OK) return st;
  if (e != hipSuccess) return qs_hip_fail(ctx, e, "qs_cf_spectrum", __FILE__, __LINE__);
  return QS_OK;
}
