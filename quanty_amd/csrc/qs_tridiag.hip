// Householder reduction of a dense Hermitian matrix to a real symmetric
// tridiagonal one, A = Q T Q^+, in one cooperative launch.
//
// The algorithm is LAPACK's unblocked zhetd2 with the lower triangle
// (reflector k from column k, pivot row k+1; v_k(k+1) = 1 and v_k(k+2:) kept in
// A(k+2:, k); T's diagonal d, off-diagonal e; tau_k), the layout the
// back-transform kernels expect. Its steps per column are
//   x = tau H22 v,  alpha = -tau/2 (x^+ v),  w = x + alpha v,
//   H22 -= v w^+ + w v^+,
// and the next reflector comes from the updated next column.
//
// Work split (one grid-wide barrier per column):
//   * columns are owned by wavefronts (column j by wave j mod the waves of the
//     grid); the matrix is Hermitian, so (H22 v)_j = column_j^+ v is a dot
//     product down an owned column, and the rank-2 update of column j touches
//     only that column: an owner updates its columns with step k's (v, w) and,
//     in the same pass, forms the next step's product with v_{k+1};
//   * the next column (k+1) is not updated by its owner: every workgroup
//     rebuilds it from memory and (v_k, w_k), and forms the reflector v_{k+1}
//     itself, so all of them hold v_{k+1} without waiting for one another;
//     fixed-order sums make the copies identical;
//   * w_k needs all of H22 v_k and the sum of the owners' partial dot
//     products, hence the one barrier per column.
// Workgroup s < nv also applies each reflector to vector s of Y (Y <- Q^+ Y),
// so the projections of the spectra come out of the same launch.
#include <algorithm>

#include "qs_internal.h"
#include "qs_kernels.h"

#define QS_HETRD_FUSED_MAX 800  // rows (kept in LDS three times: 37.5 KB at the maximum)
#ifndef QS_HETRD_BLOCK
#define QS_HETRD_BLOCK 512
#endif

namespace {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

__device__ __forceinline__ double2 block_sum2(double2 v, double *red) {
  const double r = block_sum_values(v.x, red);
  const double i = block_sum_values(v.y, red);
  return make_double2(r, i);
}

// LAPACK zlarfg on c[k0..n-1] (in LDS, the whole workgroup): alpha = c[k0],
// x = c[k0+1..]. Returns tau and beta (real); x becomes the reflector's tail
// and c[k0] its leading 1.
__device__ void block_larfg(double2 *c, int k0, int n, double2 &tau, double &beta, double *red) {
  double s = 0.0;
  for (int i = k0 + 1 + (int)threadIdx.x; i < n; i += blockDim.x) s += c[i].x * c[i].x + c[i].y * c[i].y;
  const double xn2 = block_sum_values(s, red);
  const double2 alpha = c[k0];
  if (xn2 == 0.0 && alpha.y == 0.0) {
    tau = make_double2(0.0, 0.0);
    beta = alpha.x;
  } else {
    const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
    beta = alpha.x >= 0.0 ? -nrm : nrm;
    tau = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
    // scale = 1 / (alpha - beta)
    const double dr = alpha.x - beta, di = alpha.y;
    const double dd = dr * dr + di * di;
    const double2 sc = make_double2(dr / dd, -di / dd);
    for (int i = k0 + 1 + (int)threadIdx.x; i < n; i += blockDim.x) c[i] = cmul(sc, c[i]);
  }
  __syncthreads();
  if (threadIdx.x == 0) c[k0] = make_double2(1.0, 0.0);
  __syncthreads();
}

// y <- (I - tau v v^+)^+ y = y - conj(tau) v (v^+ y), v = c[k0..n-1] (c[k0] = 1)
__device__ void block_apply_qh(const double2 *c, int k0, int n, double2 tau, double2 *y, double *red) {
  double2 s = make_double2(0.0, 0.0);
  for (int i = k0 + (int)threadIdx.x; i < n; i += blockDim.x) {
    const double2 t = cmulc(c[i], y[i]);
    s.x += t.x;
    s.y += t.y;
  }
  s = block_sum2(s, red);
  const double2 f = cmulc(tau, s);  // conj(tau) (v^+ y)
  for (int i = k0 + (int)threadIdx.x; i < n; i += blockDim.x) {
    const double2 t = cmul(c[i], f);
    y[i].x -= t.x;
    y[i].y -= t.y;
  }
  __syncthreads();
}

// Grid-wide barrier of a cooperative launch: one counter and one generation
// word (bar[0], bar[16]); workgroup leaders arrive with a device-scope atomic,
// the last one resets the counter and advances the generation, the others
// wait for it with a bounded spin. When the spin runs out, flag[0] is set and
// every workgroup leaves at its next barrier (the caller then discards the
// result). A flat barrier: on this device the cost grows with the number of
// arriving workgroups (about 70 ns each, whatever the counter layout), so
// the grids are kept small. Returns false when the workgroup must stop.
__device__ bool grid_barrier(unsigned *bar, int *flag) {
  __shared__ int stop;
  // every thread's own stores are made visible device-wide before the
  // workgroup arrives (a fence orders only the memory operations of the
  // thread that executes it)
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned *cnt = bar, *gen = bar + 16;
    const unsigned my = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (__hip_atomic_fetch_add(cnt, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1) {
      __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      long spins = 0;
      while (__hip_atomic_load(gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == my) {
        if (++spins > (1L << 26)) {
          __hip_atomic_store(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        __builtin_amdgcn_s_sleep(1);
      }
    }
    stop = __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  // and every thread sees the other workgroups' stores from here on
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return stop == 0;
}

}  // namespace

// A: n x n column-major (entry (i, j) at A[j*n + i]), overwritten as zhetrd
// (lower) leaves it. d, e, tau: n, n-1, n-1. p: 2n scratch and part:
// 2 gridDim.x scratch, each used as two buffers by step parity (the owners
// write step k+1's while slower workgroups may still read step k's). Y: nv
// vectors of n entries (Y <- Q^+ Y). Reflector k is stored into column k at the
// start of step k, when no workgroup reads that column any more.
__global__ void __launch_bounds__(QS_HETRD_BLOCK)
k_hetrd(int n, double2 *__restrict__ A, double *__restrict__ d, double *__restrict__ e, double2 *__restrict__ tau,
        double2 *__restrict__ p, double2 *__restrict__ part, int nv, double2 *__restrict__ Y, unsigned *bar,
        int *flag) {
  extern __shared__ double2 sm[];
  double2 *vk = sm;          // v_k over rows k+1..n-1
  double2 *wk = sm + n;      // w_k over rows k+1..n-1
  double2 *vn = sm + 2 * n;  // the next column, then v_{k+1}
  __shared__ double red[QS_HETRD_BLOCK / 64];
  __shared__ double2 wpart[QS_HETRD_BLOCK / 64];
  const int tid = threadIdx.x, bs = blockDim.x, lane = tid & 63, wv = tid >> 6;
  const int nwpb = bs >> 6;
  const int G = (int)gridDim.x;
  const int gw = (int)blockIdx.x * nwpb + wv, nwv = G * nwpb;
  const bool lead = blockIdx.x == 0;
  double2 *ys = (int)blockIdx.x < nv ? Y + (size_t)blockIdx.x * n : nullptr;
  auto column = [&](int j) { return A + (size_t)j * n; };

  // owners: update columns j >= r0 (rows >= r0) with (vk, wk) when `upd`, and
  // form pw[j] = column_j^+ vn when `prod`; this workgroup's sum of
  // conj(pw[j]) vn[j] goes to partw[blockIdx.x]
  auto owners = [&](int r0, bool upd, bool prod, double2 *pw, double2 *partw) {
    double2 acc = make_double2(0.0, 0.0);
    for (int j = gw; j < n; j += nwv) {
      if (j < r0) continue;
      double2 *cj = column(j);
      const double2 wj = make_double2(wk[j].x, -wk[j].y), vj = make_double2(vk[j].x, -vk[j].y);
      double2 pr = make_double2(0.0, 0.0);
      for (int i = r0 + lane; i < n; i += 64) {
        double2 a = cj[i];
        if (upd) {
          const double2 t1 = cmul(vk[i], wj), t2 = cmul(wk[i], vj);
          a.x -= t1.x + t2.x;
          a.y -= t1.y + t2.y;
          cj[i] = a;
        }
        if (prod) {
          const double2 t = cmulc(a, vn[i]);
          pr.x += t.x;
          pr.y += t.y;
        }
      }
      if (prod) {
        pr.x = wave_sum(pr.x);
        pr.y = wave_sum(pr.y);
        if (lane == 0) pw[j] = pr;
        const double2 t = cmulc(pr, vn[j]);
        acc.x += t.x;
        acc.y += t.y;
      }
    }
    if (lane == 0) wpart[wv] = acc;
    __syncthreads();
    if (tid == 0) {
      double2 s = make_double2(0.0, 0.0);
      for (int w = 0; w < nwpb; ++w) {
        s.x += wpart[w].x;
        s.y += wpart[w].y;
      }
      partw[blockIdx.x] = s;
    }
    __syncthreads();
  };

  // column 0 and reflector 0
  for (int i = tid; i < n; i += bs) vn[i] = A[i];
  __syncthreads();
  if (lead && tid == 0) d[0] = vn[0].x;
  if (n == 1) return;
  double2 tk;
  double beta;
  block_larfg(vn, 1, n, tk, beta, red);
  if (lead && tid == 0) {
    tau[0] = tk;
    e[0] = beta;
  }
  if (ys && (tk.x != 0.0 || tk.y != 0.0)) block_apply_qh(vn, 1, n, tk, ys, red);
  for (int i = 1 + tid; i < n; i += bs) vk[i] = vn[i];
  __syncthreads();
  // p_0 = H22 v_0 (columns and rows from 1) into buffer 0; vn holds v_0 too
  owners(1, false, tk.x != 0.0 || tk.y != 0.0, p, part);
  if (!grid_barrier(bar, flag)) return;

  for (int k = 0; k <= n - 2; ++k) {
    const bool act = tk.x != 0.0 || tk.y != 0.0;
    const double2 *pr = p + (size_t)(k & 1) * n, *partr = part + (size_t)(k & 1) * G;
    double2 *pw = p + (size_t)((k + 1) & 1) * n, *partw = part + (size_t)((k + 1) & 1) * G;
    // reflector k into column k (rows k+2..), beta_k at (k+1, k)
    if (lead) {
      double2 *c0 = column(k);
      for (int i = k + 2 + tid; i < n; i += bs) c0[i] = vk[i];
      if (tid == 0) c0[k + 1] = make_double2(e[k], 0.0);
    }
    // w_k = tau x + alpha v, x = H22 v_k from the owners' p, alpha = -|tau|^2/2 (p^+ v)
    if (act) {
      double2 s = make_double2(0.0, 0.0);
      for (int g = tid; g < G; g += bs) {
        s.x += partr[g].x;
        s.y += partr[g].y;
      }
      s = block_sum2(s, red);
      const double h = -0.5 * (tk.x * tk.x + tk.y * tk.y);
      const double2 al = make_double2(h * s.x, h * s.y);
      for (int i = k + 1 + tid; i < n; i += bs) {
        const double2 a = cmul(tk, pr[i]), b = cmul(al, vk[i]);
        wk[i] = make_double2(a.x + b.x, a.y + b.y);
      }
    }
    __syncthreads();
    // column k+1 (rows k+1..) after step k's update, in every workgroup
    {
      const double2 *ck = column(k + 1);
      const double2 wkk = act ? make_double2(wk[k + 1].x, -wk[k + 1].y) : make_double2(0.0, 0.0);
      const double2 vkk = act ? make_double2(vk[k + 1].x, -vk[k + 1].y) : make_double2(0.0, 0.0);
      for (int i = k + 1 + tid; i < n; i += bs) {
        double2 c = ck[i];
        if (act) {
          const double2 t1 = cmul(vk[i], wkk), t2 = cmul(wk[i], vkk);
          c.x -= t1.x + t2.x;
          c.y -= t1.y + t2.y;
        }
        vn[i] = c;
      }
    }
    __syncthreads();
    if (lead && tid == 0) d[k + 1] = vn[k + 1].x;
    if (k + 1 > n - 2) break;
    double2 tn;
    block_larfg(vn, k + 2, n, tn, beta, red);
    if (lead && tid == 0) {
      tau[k + 1] = tn;
      e[k + 1] = beta;
    }
    const bool nact = tn.x != 0.0 || tn.y != 0.0;
    if (ys && nact) block_apply_qh(vn, k + 2, n, tn, ys, red);
    // owners: step k's update of columns k+2.. and p_{k+1} = H22 v_{k+1}
    owners(k + 2, act, nact, pw, partw);
    for (int i = k + 2 + tid; i < n; i += bs) vk[i] = vn[i];
    tk = tn;
    if (!grid_barrier(bar, flag)) return;
  }
}

// In-place tridiagonal reduction of the n x n Hermitian matrix A (device,
// column-major, lower triangle as zhetrd uses it) with Y <- Q^+ Y for nv
// vectors: fills d, e, tau (device). Returns QS_EUNSUPPORTED, leaving A and Y
// untouched, above QS_HETRD_FUSED_MAX rows or when the device cannot hold the
// cooperative grid; the caller then takes another route. Measured on the
// MI355X (eigenvalues and projections of 3 vectors for a dense complex matrix,
// host QL included): 1.5 against 2.0 ms for rocSOLVER's zhetrd at n = 120,
// 2.8 against 4.3 ms at 190, 10.6 against 13.8 ms at 500, a tie at 800, and
// 66 against 58 ms at 1410: every column costs one barrier and a few
// dependent global round trips, so the launch count stops mattering once
// columns are long.
qs_status qs_hetrd_fused(qs_ctx *ctx, int n, double2 *A, double *d, double *e, double2 *tau, int nv, double2 *Y) {
  if (n > QS_HETRD_FUSED_MAX || n < 1) return QS_EUNSUPPORTED;
  const size_t shm = (size_t)3 * n * sizeof(double2);
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_hetrd, QS_HETRD_BLOCK, shm) != hipSuccess ||
      per_cu < 1) {
    (void)hipGetLastError();
    return QS_EUNSUPPORTED;
  }
  // few workgroups: a barrier costs about 70 ns per arriving workgroup; 4..16
  // were the fastest grids from n = 120 to 500
  int G = std::max(4, std::min(16, n / 16));
  G = std::min(G, ctx->num_cu);
  G = std::max(G, std::min(nv, ctx->num_cu));
  if (nv > G) return QS_EUNSUPPORTED;
  void *scr;
  const size_t pbytes = 2 * ((size_t)n + (size_t)G) * sizeof(double2);
  QS_TRY(qs_scratch(ctx, pbytes + 256, &scr));
  double2 *p = (double2 *)scr;
  double2 *part = p + 2 * (size_t)n;
  unsigned *bar = (unsigned *)((char *)scr + pbytes);  // counter, generation, flag
  int *flag = (int *)(bar + 32);
  QS_HIP(ctx, hipMemsetAsync(bar, 0, 256, ctx->stream));
  void *args[] = {&n, &A, &d, &e, &tau, &p, &part, &nv, &Y, &bar, &flag};
  const hipError_t err = hipLaunchCooperativeKernel((const void *)k_hetrd, dim3(G), dim3(QS_HETRD_BLOCK), args, shm,
                                                    ctx->stream);
  if (err == hipErrorCooperativeLaunchTooLarge || err == hipErrorNotSupported) {
    (void)hipGetLastError();
    return QS_EUNSUPPORTED;
  }
  QS_HIP(ctx, err);
  int hflag = 0;
  QS_TRY(qs_d2h(ctx, &hflag, flag, sizeof(int)));
  if (hflag) return qs_fail(ctx, QS_EHIP, "dense reduction: a grid barrier timed out");
  return QS_OK;
}
