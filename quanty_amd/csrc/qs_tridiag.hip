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
// Work split:
//   * columns are owned by workgroups (column j by workgroup j mod G) and held
//     in registers for the whole launch; the matrix is Hermitian, so
//     (H22 v)_j = column_j^+ v is a dot product down an owned column, and the
//     rank-2 update of column j touches only that column: an owner updates its
//     columns with step k's (v, w) and, in the same pass, forms the next
//     step's product with v_{k+1};
//   * the next column (k+1) is not updated by its owner: every workgroup
//     rebuilds it from the copy its owner published and (v_k, w_k), and forms
//     the reflector v_{k+1} itself, so all of them hold v_{k+1} without waiting
//     for one another; fixed-order sums make the copies identical;
//   * w_k needs all of H22 v_k and the sum of the owners' partial dot
//     products: the one exchange between workgroups per column.
// Workgroup s < nv also applies each reflector to vector s of Y (Y <- Q^+ Y),
// so the projections of the spectra come out of the same launch.
#include <algorithm>

#include "qs_internal.h"
#include "qs_kernels.h"

namespace {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// ---------------------------------------------------------------------------
// Layout and exchange. Up to 256 workgroups of 512 threads (one per CU, two waves per SIMD, so a
// thread has 256 VGPRs); workgroup g owns columns g, g + G, ... (at most
// QS_HETRD_REG_NC of them) and thread t holds rows t, t + 512, t + 1024,
// t + 1536 of each owned column for the whole launch, so the matrix is read
// once and never written back.
//
// What workgroups exchange per column (the owners' entries of H22 v, one
// partial dot product per workgroup, the next column as its owner left it)
// goes through device memory with device-scope loads and stores, which are
// coherent across the XCDs' L2 caches without writing them back or
// invalidating them. There is no barrier: every 8-byte word of the exchange
// buffers starts as a marker (all bits set, a NaN no arithmetic produces) and
// a reader polls the words it needs until none is the marker. The partial
// dot products, one per workgroup and written every column, make the wait a
// barrier: a workgroup that holds all partials of step k knows that every
// workgroup has finished reading the buffers of step k-1. The buffers rotate
// over three copies (step k reads copy k mod 3, writes copy k+1 mod 3), and in
// step k a workgroup puts the marker back into its own words of copy
// k+2 mod 3, read last in step k-1; those stores complete before it writes
// anything of step k+1's, so a reader of copy k+2 mod 3 sees the marker or
// the new value, never the old one. Compared with per-workgroup flags (store,
// wait for completion, set the flag, poll it, then load the data: about four
// dependent round trips to memory per column) this is one round trip.
//
// Between writing its step k+1 data and polling for it, a workgroup does the
// work nobody waits for: storing reflector k into A and applying reflector
// k+1 to its vector of Y.
#define QS_HETRD_REG_BLOCK 512
#define QS_HETRD_REG_NC 6   // owned columns per workgroup
#define QS_HETRD_REG_RPT 4  // rows per thread
#define QS_HETRD_REG_MAX (QS_HETRD_REG_BLOCK * QS_HETRD_REG_RPT)

constexpr unsigned long long QS_MARK = ~0ull;

__device__ __forceinline__ void st_word(double *p, double v) {
  __hip_atomic_store(reinterpret_cast<unsigned long long *>(p), (unsigned long long)__double_as_longlong(v),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_dev(double2 *p, double2 v) {
  st_word(&p->x, v.x);
  st_word(&p->y, v.y);
}

__device__ __forceinline__ void st_mark(double2 *p) {
  unsigned long long *q = reinterpret_cast<unsigned long long *>(p);
  __hip_atomic_store(q, QS_MARK, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(q + 1, QS_MARK, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Up to NP exchanged values at once: every load is issued before any is
// waited for, and only the words still holding the marker are loaded again.
// Bounded: on running out, or when another workgroup has, err is raised and
// false returned. use[q] = false skips value q.
template <int NP>
__device__ __forceinline__ bool ld_wait_n(const double2 *const (&p)[NP], const bool (&use)[NP], double2 (&v)[NP],
                                          int *err) {
  unsigned long long w[2 * NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    const unsigned long long *a = reinterpret_cast<const unsigned long long *>(p[q]);
    w[2 * q] = use[q] ? __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0ull;
    w[2 * q + 1] = use[q] ? __hip_atomic_load(a + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0ull;
  }
  long spins = 0;
  for (;;) {
    bool all = true;
#pragma unroll
    for (int q = 0; q < 2 * NP; ++q) all = all && w[q] != QS_MARK;
    if (all) break;
    if (++spins > (1L << 22)) {
      __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
    __builtin_amdgcn_s_sleep(1);
#pragma unroll
    for (int q = 0; q < 2 * NP; ++q)
      if (w[q] == QS_MARK)
        w[q] = __hip_atomic_load(reinterpret_cast<const unsigned long long *>(p[q / 2]) + (q & 1), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
  }
#pragma unroll
  for (int q = 0; q < NP; ++q)
    v[q] = make_double2(__longlong_as_double((long long)w[2 * q]), __longlong_as_double((long long)w[2 * q + 1]));
  return true;
}

// Wave sums of 8 values per lane by recursive halving: at lane distances 32,
// 16 and 8 each lane keeps half of its values and adds its partner's copy of
// them, then the last value is summed over lane distances 4, 2 and 1. Returns
// the wave sum of value 4 b5 + 2 b4 + b3 (lane bits b5, b4, b3); 10 shuffles
// instead of 48, in a fixed order.
__device__ __forceinline__ double wave_sum8(const double (&v)[8], int lane) {
  double w4[4], w2[2];
  {
    const bool hi = lane & 32;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const double send = hi ? v[i] : v[4 + i];
      const double keep = hi ? v[4 + i] : v[i];
      w4[i] = keep + __shfl_xor(send, 32);
    }
  }
  {
    const bool hi = lane & 16;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const double send = hi ? w4[i] : w4[2 + i];
      const double keep = hi ? w4[2 + i] : w4[i];
      w2[i] = keep + __shfl_xor(send, 16);
    }
  }
  double x;
  {
    const bool hi = lane & 8;
    const double send = hi ? w2[0] : w2[1];
    const double keep = hi ? w2[1] : w2[0];
    x = keep + __shfl_xor(send, 8);
  }
  x += __shfl_xor(x, 4);
  x += __shfl_xor(x, 2);
  x += __shfl_xor(x, 1);
  return x;
}

// block sum of one complex value per thread, fixed order, valid in every
// thread; red: blockDim.x / 32 doubles
__device__ __forceinline__ double2 block_sum_c(double2 v, double *red) {
  v.x = wave_sum(v.x);
  v.y = wave_sum(v.y);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[2 * w] = v.x;
    red[2 * w + 1] = v.y;
  }
  __syncthreads();
  double2 r = make_double2(0.0, 0.0);
  for (int i = 0; i < nw; ++i) {
    r.x += red[2 * i];
    r.y += red[2 * i + 1];
  }
  __syncthreads();
  return r;
}

// LAPACK zlarfg on c[k0..n-1] (in LDS, the whole workgroup): alpha = c[k0],
// x = c[k0+1..]. Returns tau and beta (real); x becomes the reflector's tail
// and c[k0] its leading 1.
__device__ void reg_larfg(double2 *c, int k0, int n, double2 &tau, double &beta, double *red) {
  double s = 0.0;
  for (int i = k0 + 1 + (int)threadIdx.x; i < n; i += blockDim.x) s += c[i].x * c[i].x + c[i].y * c[i].y;
  const double xn2 = block_sum_c(make_double2(s, 0.0), red).x;
  const double2 alpha = c[k0];
  if (xn2 == 0.0 && alpha.y == 0.0) {
    tau = make_double2(0.0, 0.0);
    beta = alpha.x;
  } else {
    const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
    beta = alpha.x >= 0.0 ? -nrm : nrm;
    tau = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
    const double dr = alpha.x - beta, di = alpha.y;
    const double dd = dr * dr + di * di;
    const double2 sc = make_double2(dr / dd, -di / dd);
    for (int i = k0 + 1 + (int)threadIdx.x; i < n; i += blockDim.x) c[i] = cmul(sc, c[i]);
  }
  __syncthreads();
  if (threadIdx.x == 0) c[k0] = make_double2(1.0, 0.0);
  __syncthreads();
}

// y <- (I - tau v v^+)^+ y, v = c[k0..n-1] (c[k0] = 1), one complex reduction
__device__ void reg_apply_qh(const double2 *c, int k0, int n, double2 tau, double2 *y, double *red) {
  double2 s = make_double2(0.0, 0.0);
  for (int i = k0 + (int)threadIdx.x; i < n; i += blockDim.x) {
    const double2 t = cmulc(c[i], y[i]);
    s.x += t.x;
    s.y += t.y;
  }
  s = block_sum_c(s, red);
  const double2 f = cmulc(tau, s);
  for (int i = k0 + (int)threadIdx.x; i < n; i += blockDim.x) {
    const double2 t = cmul(c[i], f);
    y[i].x -= t.x;
    y[i].y -= t.y;
  }
  __syncthreads();
}

}  // namespace

// A: n x n column-major, lower triangle read (the upper one is taken as its
// conjugate transpose), overwritten as zhetrd (lower) leaves it. pbuf, colbuf:
// 3n each, part: 3G, every word the marker at launch. err: zero at launch.
__global__ void __launch_bounds__(QS_HETRD_REG_BLOCK)
k_hetrd_reg(int n, double2 *__restrict__ A, double *__restrict__ d, double *__restrict__ e,
            double2 *__restrict__ tau, double2 *pbuf, double2 *part, double2 *colbuf, int nv,
            double2 *__restrict__ Y, int *err) {
  extern __shared__ double2 sm[];
  double2 *vk = sm, *wk = sm + n, *vn = sm + 2 * n;
  __shared__ double red[2 * (QS_HETRD_REG_BLOCK / 64)];
  __shared__ double colsum[QS_HETRD_REG_BLOCK / 64][16];
  __shared__ double2 pcon[QS_HETRD_REG_NC];
  __shared__ double2 scal;
  __shared__ int bad;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, bs = QS_HETRD_REG_BLOCK;
  const int G = (int)gridDim.x, g = (int)blockIdx.x;
  const bool lead = g == 0;
  double2 *ys = g < nv ? Y + (size_t)g * n : nullptr;

  // owned columns into registers (rows below the diagonal from the column,
  // rows above it from the conjugated row of the lower triangle)
  double2 a[QS_HETRD_REG_NC][QS_HETRD_REG_RPT];
#pragma unroll
  for (int c = 0; c < QS_HETRD_REG_NC; ++c) {
    const int j = g + c * G;
#pragma unroll
    for (int r = 0; r < QS_HETRD_REG_RPT; ++r) {
      const int i = tid + r * bs;
      double2 v = make_double2(0.0, 0.0);
      if (j < n && i < n) {
        if (i >= j) {
          v = A[(size_t)j * n + i];
        } else {
          v = A[(size_t)i * n + j];
          v.y = -v.y;
        }
      }
      a[c][r] = v;
    }
  }

  // owners: update owned columns j >= r0 (rows >= r0) with (vk, wk) when
  // `upd`; the owner of column r0 stores it (rows >= r0) to colw; with `prod`,
  // p[j] = column_j^+ vn goes to pw; this workgroup's sum of conj(p[j]) vn[j]
  // (zero without `prod`) goes to partw[g] last.
  auto owners = [&](int r0, bool upd, bool prod, double2 *pw, double2 *partw, double2 *colw) {
    static_assert(QS_HETRD_REG_NC == 6, "two groups of three columns");
    // two groups of three owned columns; in each, rows outer (a row's v, w and
    // next v read once per group) and the group's columns inner
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      double acc[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) acc[q] = 0.0;
#pragma unroll
      for (int r = 0; r < QS_HETRD_REG_RPT; ++r) {
        const int i = tid + r * bs;
        if (i < r0 || i >= n) continue;
        const double2 vi = upd ? vk[i] : make_double2(0.0, 0.0);
        const double2 wi = upd ? wk[i] : make_double2(0.0, 0.0);
        const double2 ni = prod ? vn[i] : make_double2(0.0, 0.0);
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
          const int c = 3 * h + cc;
          const int j = g + c * G;
          if (j < r0 || j >= n) continue;
          double2 x = a[c][r];
          if (upd) {
            const double2 wj = make_double2(wk[j].x, -wk[j].y), vj = make_double2(vk[j].x, -vk[j].y);
            const double2 t1 = cmul(vi, wj), t2 = cmul(wi, vj);
            x.x -= t1.x + t2.x;
            x.y -= t1.y + t2.y;
            a[c][r] = x;
          }
          if (prod) {
            const double2 t = cmulc(x, ni);
            acc[2 * cc] += t.x;
            acc[2 * cc + 1] += t.y;
          }
          if (j == r0) st_dev(&colw[i], x);
        }
      }
      if (prod) {
        const double x = wave_sum8(acc, lane);
        if ((lane & 7) == 0) colsum[wv][8 * h + (lane >> 3)] = x;  // value 4 b5 + 2 b4 + b3 of group h

      }
    }
    if (prod) {
      __syncthreads();
      if (tid < QS_HETRD_REG_NC) {
        const int j = g + tid * G;
        double2 s = make_double2(0.0, 0.0);
        for (int w = 0; w < bs / 64; ++w) {
          const int q = 8 * (tid / 3) + 2 * (tid % 3);  // column tid: value 2 (tid mod 3) of group tid / 3
          s.x += colsum[w][q];
          s.y += colsum[w][q + 1];
        }
        double2 con = make_double2(0.0, 0.0);
        if (j >= r0 && j < n) {
          st_dev(&pw[j], s);
          con = cmulc(s, vn[j]);
        }
        pcon[tid] = con;
      }
      __syncthreads();
    }
    if (tid == 0) {
      double2 s = make_double2(0.0, 0.0);
      if (prod)
        for (int c = 0; c < QS_HETRD_REG_NC; ++c) {
          s.x += pcon[c].x;
          s.y += pcon[c].y;
        }
      st_dev(&partw[g], s);
    }
  };

  // column 0 and reflector 0
  for (int i = tid; i < n; i += bs) vn[i] = A[i];
  __syncthreads();
  if (lead && tid == 0) d[0] = vn[0].x;
  if (n == 1) return;
  double2 tk;
  double beta;
  reg_larfg(vn, 1, n, tk, beta, red);
  if (lead && tid == 0) {
    tau[0] = tk;
    e[0] = beta;
  }
  // p_0 = H22 v_0, column 1 as it is and the partials: copy 0
  owners(1, false, tk.x != 0.0 || tk.y != 0.0, pbuf, part, colbuf);
  if (ys && (tk.x != 0.0 || tk.y != 0.0)) reg_apply_qh(vn, 1, n, tk, ys, red);
  {
    double2 *t = vk;
    vk = vn;
    vn = t;
  }

  for (int k = 0; k <= n - 2; ++k) {
    const bool act = tk.x != 0.0 || tk.y != 0.0;
    const int cr = k % 3, cw = (k + 1) % 3, cz = (k + 2) % 3;
    const double2 *pr = pbuf + (size_t)cr * n, *partr = part + (size_t)cr * G, *colr = colbuf + (size_t)cr * n;
    // step k's data: the partials of every workgroup (wave 0), this thread's
    // rows of p_k and of column k+1
    if (tid == 0) bad = 0;
    __syncthreads();
    bool ok = true;
    double2 s = make_double2(0.0, 0.0);
    if (wv == 0) {
      // the partials: G <= 256, four per lane
      const double2 *pp[4];
      bool use[4];
      double2 v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        use[q] = lane + 64 * q < G;
        pp[q] = partr + (use[q] ? lane + 64 * q : 0);
      }
      ok = ld_wait_n<4>(pp, use, v, err);
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (use[q]) {
          s.x += v[q].x;
          s.y += v[q].y;
        }
    }
    {
      // this thread's rows of column k+1 and of p_k
      const double2 *pp[2 * QS_HETRD_REG_RPT];
      bool use[2 * QS_HETRD_REG_RPT];
      double2 v[2 * QS_HETRD_REG_RPT];
#pragma unroll
      for (int r = 0; r < QS_HETRD_REG_RPT; ++r) {
        const int i = k + 1 + tid + r * bs;
        use[2 * r] = i < n;
        use[2 * r + 1] = i < n && act;
        pp[2 * r] = colr + (i < n ? i : 0);
        pp[2 * r + 1] = pr + (i < n ? i : 0);
      }
      ok = ld_wait_n<2 * QS_HETRD_REG_RPT>(pp, use, v, err) && ok;
#pragma unroll
      for (int r = 0; r < QS_HETRD_REG_RPT; ++r) {
        const int i = k + 1 + tid + r * bs;
        if (i < n) {
          vn[i] = v[2 * r];
          if (act) wk[i] = v[2 * r + 1];
        }
      }
    }
    if (!ok) bad = 1;
    if (wv == 0) {
      s.x = wave_sum(s.x);
      s.y = wave_sum(s.y);
      if (lane == 0) scal = s;
    }
    __syncthreads();
    if (bad) return;
    // every workgroup has finished step k-1: the marker back into this
    // workgroup's words of copy k+2 mod 3 (its partial, its p entries, and
    // column k+3 if it owns it)
    {
      double2 *pz = pbuf + (size_t)cz * n, *colz = colbuf + (size_t)cz * n;
      if (tid == 0) st_mark(&part[(size_t)cz * G + g]);
      if (tid < QS_HETRD_REG_NC && g + tid * G < n) st_mark(&pz[g + tid * G]);
      if (k + 3 < n && (k + 3) % G == g)
        for (int i = k + 3 + tid; i < n; i += bs) st_mark(&colz[i]);
    }
    // w_k = tau p + alpha v, alpha = -|tau|^2/2 (p^+ v)
    if (act) {
      const double h = -0.5 * (tk.x * tk.x + tk.y * tk.y);
      const double2 al = make_double2(h * scal.x, h * scal.y);
      for (int i = k + 1 + tid; i < n; i += bs) {
        const double2 x = cmul(tk, wk[i]), y = cmul(al, vk[i]);
        wk[i] = make_double2(x.x + y.x, x.y + y.y);
      }
    }
    __syncthreads();
    // column k+1 after step k's update
    if (act) {
      const double2 wkk = make_double2(wk[k + 1].x, -wk[k + 1].y);
      const double2 vkk = make_double2(vk[k + 1].x, -vk[k + 1].y);
      for (int i = k + 1 + tid; i < n; i += bs) {
        double2 c = vn[i];
        const double2 t1 = cmul(vk[i], wkk), t2 = cmul(wk[i], vkk);
        c.x -= t1.x + t2.x;
        c.y -= t1.y + t2.y;
        vn[i] = c;
      }
    }
    __syncthreads();
    if (lead && tid == 0) d[k + 1] = vn[k + 1].x;
    if (k + 1 > n - 2) {
      if (lead && tid == 0) A[(size_t)k * n + k + 1] = make_double2(e[k], 0.0);
      break;
    }
    double2 tn;
    reg_larfg(vn, k + 2, n, tn, beta, red);
    if (lead && tid == 0) {
      tau[k + 1] = tn;
      e[k + 1] = beta;
    }
    const bool nact = tn.x != 0.0 || tn.y != 0.0;
    // every thread's markers have completed before anything of step k+1 is
    // stored (by whichever thread)
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    owners(k + 2, act, nact, pbuf + (size_t)cw * n, part + (size_t)cw * G, colbuf + (size_t)cw * n);
    // off the critical path: reflector k into column k, reflector k+1 on Y
    if (lead) {
      double2 *c0 = A + (size_t)k * n;
      for (int i = k + 2 + tid; i < n; i += bs) c0[i] = vk[i];
      if (tid == 0) c0[k + 1] = make_double2(e[k], 0.0);
    }
    if (ys && nact) reg_apply_qh(vn, k + 2, n, tn, ys, red);
    __syncthreads();
    {
      double2 *t = vk;
      vk = vn;
      vn = t;
    }
    tk = tn;
  }
}

// In-place tridiagonal reduction of the n x n Hermitian matrix A (device,
// column-major, lower triangle read as zhetrd reads it) with Y <- Q^+ Y for nv
// vectors: fills d, e, tau (device). Returns QS_EUNSUPPORTED, leaving A and Y
// untouched, when the grid (one workgroup per QS_HETRD_REG_NC columns, at
// least nv, at most 256) does not fit the device, i.e. above 1536 rows on the
// MI355X; the caller then takes another route. Measured on the MI355X
// (eigenvalues and projections of 3 vectors for a random dense complex matrix,
// tridiagonal eigenproblem included): 1.15 against 2.0 ms with rocSOLVER's
// zhetrd at n = 120, 3.8 against 7.9 ms at 300, 10.0 against 20.3 ms at 800,
// 19.3 against 38.7 ms at 1410. Per column (n = 1410, timestamps inside the
// kernel): about 9 us, of which 3.4 waiting for the exchanged data (memory
// latency), 2.7 the owners' update and products, 1.5 the reflector.
qs_status qs_hetrd_fused(qs_ctx *ctx, int n, double2 *A, double *d, double *e, double2 *tau, int nv, double2 *Y) {
  if (n < 1 || n > QS_HETRD_REG_MAX) return QS_EUNSUPPORTED;
  const size_t shm = (size_t)3 * n * sizeof(double2);
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_hetrd_reg, QS_HETRD_REG_BLOCK, shm) != hipSuccess ||
      per_cu < 1) {
    (void)hipGetLastError();
    return QS_EUNSUPPORTED;
  }
  int G = (n + QS_HETRD_REG_NC - 1) / QS_HETRD_REG_NC;
  G = std::max(G, nv);
  // the partials are polled four per lane of one wave
  if (G > per_cu * ctx->num_cu || G > 256) return QS_EUNSUPPORTED;
  void *scr;
  const size_t xbytes = 3 * (2 * (size_t)n + (size_t)G) * sizeof(double2);
  QS_TRY(qs_scratch(ctx, xbytes + 64, &scr));
  double2 *pbuf = (double2 *)scr;
  double2 *colbuf = pbuf + 3 * (size_t)n;
  double2 *part = colbuf + 3 * (size_t)n;
  int *err = (int *)((char *)scr + xbytes);
  // every exchange word starts as the marker (all bits set), err as zero
  QS_HIP(ctx, hipMemsetAsync(scr, 0xff, xbytes, ctx->stream));
  QS_HIP(ctx, hipMemsetAsync(err, 0, 64, ctx->stream));
  void *args[] = {&n, &A, &d, &e, &tau, &pbuf, &part, &colbuf, &nv, &Y, &err};
  const hipError_t herr = hipLaunchCooperativeKernel((const void *)k_hetrd_reg, dim3(G), dim3(QS_HETRD_REG_BLOCK),
                                                     args, shm, ctx->stream);
  if (herr == hipErrorCooperativeLaunchTooLarge || herr == hipErrorNotSupported) {
    (void)hipGetLastError();
    return QS_EUNSUPPORTED;
  }
  QS_HIP(ctx, herr);
  int hflag = 0;
  QS_TRY(qs_d2h(ctx, &hflag, err, sizeof(int)));
  if (hflag) return qs_fail(ctx, QS_EHIP, "dense reduction: a workgroup waited too long for its data");
  return QS_OK;
}
