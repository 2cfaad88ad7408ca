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
// Workgroup s < nv also applies each reflector to vector s of Y (Y <- Q^+ Y,
// the vector in LDS), so the projections of the spectra come out of the same
// launch.
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
// QS_HETRD_REG_NC of them) and thread t holds rows t, t + 512, t + 1024 of
// each owned column for the whole launch, so the matrix is read once and
// never written back.
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
// over three copies (step k reads copy k mod 3, writes copy k+1 mod 3), and at
// the end of step k a workgroup puts the marker back into its own words of
// copy k+2 mod 3, read last in step k-1; those stores complete before it
// writes anything of step k+1's, so a reader of copy k+2 mod 3 sees the marker
// or the new value, never the old one. Compared with per-workgroup flags
// (store, wait for completion, set the flag, poll it, then load the data:
// about four dependent round trips to memory per column) this is one round
// trip.
//
// Inside a workgroup a column takes three barriers (the exchanged data in, the
// norm of the next column and v^+ y, the next reflector in LDS) and one more
// in the owners' pass; wave sums go through lane exchanges (permlane swaps,
// DPP) rather than LDS. Between writing its step k+1 data and polling for it,
// the lead workgroup stores reflector k into A, which nobody waits for.
#define QS_HETRD_REG_BLOCK 512
#define QS_HETRD_REG_NC 6   // owned columns per workgroup
#define QS_HETRD_REG_RPT 3  // rows per thread
#define QS_HETRD_REG_MAX (QS_HETRD_REG_BLOCK * QS_HETRD_REG_RPT)  // = 256 workgroups x 6 columns

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
                                          int *err, long max_spins) {
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
    if (++spins > max_spins) {
      __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    if ((spins & 63) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
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

// Lane exchanges without LDS (gfx950 permlane swaps), on the two halves of a
// double. fold32(a, b): a(l) + a(l+32) in lanes l < 32, b(l-32) + b(l) in
// lanes l >= 32; fold16 likewise between rows 2r and 2r+1 of 16 lanes.
struct double4_ {
  double v[4];
};
__device__ __forceinline__ double fold32(double a, double b) {
  const unsigned long long ua = (unsigned long long)__double_as_longlong(a);
  const unsigned long long ub = (unsigned long long)__double_as_longlong(b);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)ua, (unsigned)ub, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
  return join_f64(lo[0], hi[0]) + join_f64(lo[1], hi[1]);
}
__device__ __forceinline__ double fold16(double a, double b) {
  const unsigned long long ua = (unsigned long long)__double_as_longlong(a);
  const unsigned long long ub = (unsigned long long)__double_as_longlong(b);
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)ua, (unsigned)ub, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
  return join_f64(lo[0], hi[0]) + join_f64(lo[1], hi[1]);
}
// Wave sums of 8 values per lane by recursive halving: at lane distances 32
// and 16 each lane keeps half of its values and adds its partner's copy of
// them, then pairs lanes l and 15-l for one value and sums over 8 lanes.
// Returns the wave sum of value 4 b5 + 2 b4 + b3 (lane bits b5, b4, b3), in a
// fixed order; whole wave active.
__device__ __forceinline__ double wave_sum8(const double (&v)[8], int lane) {
  double w4[4], w2[2];
#pragma unroll
  for (int i = 0; i < 4; ++i) w4[i] = fold32(v[i], v[4 + i]);
#pragma unroll
  for (int i = 0; i < 2; ++i) w2[i] = fold16(w4[i], w4[2 + i]);
  const bool hi = lane & 8;
  const double send = hi ? w2[0] : w2[1];
  const double keep = hi ? w2[1] : w2[0];
  return sum8_lanes(keep + dpp_f64<0x140>(send));
}

// Wave sums of 4 values per lane: the sum of value 2 b5 + b4 in every lane of
// row b5 b4 (lanes 16 (2 b5 + b4) onwards); fixed order, whole wave active.
__device__ __forceinline__ double wave_sum4(const double (&v)[4]) {
  const double w0 = fold32(v[0], v[2]), w1 = fold32(v[1], v[3]);
  double x = fold16(w0, w1);
  x = sum8_lanes(x);
  return x + dpp_f64<0x140>(x);
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

}  // namespace

// A: n x n column-major, lower triangle read (the upper one is taken as its
// conjugate transpose), overwritten as zhetrd (lower) leaves it. pbuf, colbuf:
// 3n each, part: 3G, every word the marker at launch. err: zero at launch.
__global__ void __launch_bounds__(QS_HETRD_REG_BLOCK)
k_hetrd_reg(int n, double2 *__restrict__ A, double *__restrict__ d, double *__restrict__ e,
            double2 *__restrict__ tau, double2 *pbuf, double2 *part, double2 *colbuf, int nv,
            double2 *__restrict__ Y, int *err, long max_spins) {
  extern __shared__ double2 sm[];
  double2 *vk = sm, *wk = sm + n, *vn = sm + 2 * n;
  __shared__ double red[4 * (QS_HETRD_REG_BLOCK / 64)];
  __shared__ __align__(16) double colsum[QS_HETRD_REG_BLOCK / 64][16];
  __shared__ double2 alph;  // row k+2 of column k+1 (alpha of the next reflector)
  __shared__ double2 shx[2];  // p^+ v over all workgroups, p at row k+1
  __shared__ int shok;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, bs = QS_HETRD_REG_BLOCK;
  const int G = (int)gridDim.x, g = (int)blockIdx.x;
  const bool lead = g == 0;
  double2 *ys = g < nv ? Y + (size_t)g * n : nullptr;
  // workgroup g < nv: vector g of Y in LDS, each row read and written only by
  // the thread that holds it (rows as the owned columns')
  double2 *yr = sm + 3 * n;
  if (ys)
    for (int i = tid; i < n; i += bs) yr[i] = ys[i];

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
      // the group's columns are g + (3h + c) G, c = 0, 1, 2: none in [r0, n)
      // when the last is below r0 or the first at n or beyond (the same for
      // the whole workgroup, so the wave sums below stay whole-wave)
      if (g + (3 * h + 2) * G < r0 || g + 3 * h * G >= n) continue;
      // the group's (conjugated) entries of w_k and v_k, read once
      double2 wj[3], vj[3];
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        const int j = min(g + (3 * h + cc) * G, n - 1);
        wj[cc] = wk[j];
        vj[cc] = vk[j];
        wj[cc].y = -wj[cc].y;
        vj[cc].y = -vj[cc].y;
      }
      double acc[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) acc[q] = 0.0;
#pragma unroll
      for (int r = 0; r < QS_HETRD_REG_RPT; ++r) {
        const int i = tid + r * bs;
        // read whether used or not (in one batch; the index kept in range)
        const int ic = min(i, n - 1);
        const double2 vi = vk[ic], wi = wk[ic], ni = vn[ic];
        if (i < r0 || i >= n) continue;
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
          const int c = 3 * h + cc;
          const int j = g + c * G;
          if (j < r0 || j >= n) continue;
          double2 x = a[c][r];
          if (upd) {
            const double2 t1 = cmul(vi, wj[cc]), t2 = cmul(wi, vj[cc]);
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
      if (wv == 0) {
        // lane 8 c + w (c < 6): wave w's sum for owned column c, all read at
        // once; the 8 lanes of column c add them (DPP), lane 8 c stores p_j
        // and forms conj(p_j) v_j; the six terms added in column order from
        // scalar reads of those lanes (within wave 0, no barrier)
        const int c = lane >> 3, w = lane & 7;
        double2 sv = make_double2(0.0, 0.0);
        if (c < QS_HETRD_REG_NC) sv = *reinterpret_cast<const double2 *>(&colsum[w][8 * (c / 3) + 2 * (c % 3)]);
        sv.x = sum8_lanes(sv.x);
        sv.y = sum8_lanes(sv.y);
        double2 con = make_double2(0.0, 0.0);
        const int j = g + c * G;
        if (w == 0 && c < QS_HETRD_REG_NC && j >= r0 && j < n) {
          st_dev(&pw[j], sv);
          con = cmulc(sv, vn[j]);
        }
        double2 tot = make_double2(0.0, 0.0);
#pragma unroll
        for (int cc = 0; cc < QS_HETRD_REG_NC; ++cc) {
          tot.x += readlane_f64(con.x, 8 * cc);
          tot.y += readlane_f64(con.y, 8 * cc);
        }
        if (lane == 0) st_dev(&partw[g], tot);
      }
    } else if (tid == 0) {
      st_dev(&partw[g], make_double2(0.0, 0.0));
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
  double ek = beta;  // e[k] of the loop's step k
  {
    double2 *t = vk;
    vk = vn;
    vn = t;
  }

  for (int k = 0; k <= n - 2; ++k) {
    const bool act = tk.x != 0.0 || tk.y != 0.0;
    const int cr = k % 3, cw = (k + 1) % 3, cz = (k + 2) % 3;
    const double2 *pr = pbuf + (size_t)cr * n, *partr = part + (size_t)cr * G, *colr = colbuf + (size_t)cr * n;
    // step k's data: the partials of every workgroup and p_k = H22 v_k at row
    // k+1 (wave 0, into LDS), this thread's rows of column k+1 and of p_k
    bool ok = true;
    if (wv == 0) {
      const double2 *pp[5];
      bool use[5];
      double2 v[5];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        use[q] = lane + 64 * q < G;
        pp[q] = partr + (use[q] ? lane + 64 * q : 0);
      }
      use[4] = act && lane == 0;
      pp[4] = pr + k + 1;
      ok = ld_wait_n<5>(pp, use, v, err, max_spins);
      double2 s = make_double2(0.0, 0.0);
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (use[q]) {
          s.x += v[q].x;
          s.y += v[q].y;
        }
      const double4_ q = {s.x, s.y, 0.0, 0.0};
      const double t = wave_sum4(q.v);  // s.x in lanes 0.., s.y in lanes 16..
      const bool all = __all(ok);
      if (lane == 0) {
        shx[0].x = t;
        shx[1] = act ? v[4] : make_double2(0.0, 0.0);
        shok = all;
      }
      if (lane == 16) shx[0].y = t;
    }
    double2 cv[QS_HETRD_REG_RPT], pv[QS_HETRD_REG_RPT];
    {
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
      ok = ld_wait_n<2 * QS_HETRD_REG_RPT>(pp, use, v, err, max_spins) && ok;
#pragma unroll
      for (int r = 0; r < QS_HETRD_REG_RPT; ++r) {
        cv[r] = v[2 * r];
        pv[r] = v[2 * r + 1];
      }
    }
    __syncthreads();
    const double2 s = shx[0], pk1 = shx[1];
    ok = ok && shok;
    // one pass over this thread's rows, held in registers: w_k = tau p +
    // alpha v (alpha = -|tau|^2/2 p^+ v, w_k at row k+1 formed by every
    // thread), column k+1 after step k's update, and the squares of its
    // entries below row k+2 (the next reflector's norm)
    double nrm = 0.0;
    {
      const double h = -0.5 * (tk.x * tk.x + tk.y * tk.y);
      const double2 al = make_double2(h * s.x, h * s.y);
      const double2 vkk = vk[k + 1];
      const double2 wk1 = act ? make_double2(tk.x * pk1.x - tk.y * pk1.y + al.x * vkk.x - al.y * vkk.y,
                                             tk.x * pk1.y + tk.y * pk1.x + al.x * vkk.y + al.y * vkk.x)
                              : make_double2(0.0, 0.0);
      const double2 wkc = make_double2(wk1.x, -wk1.y), vkc = make_double2(vkk.x, -vkk.y);
      double2 vr[QS_HETRD_REG_RPT];  // read in one batch, the index kept in range
#pragma unroll
      for (int r = 0; r < QS_HETRD_REG_RPT; ++r) vr[r] = vk[min(k + 1 + tid + r * bs, n - 1)];
#pragma unroll
      for (int r = 0; r < QS_HETRD_REG_RPT; ++r) {
        const int i = k + 1 + tid + r * bs;
        if (i >= n) continue;
        double2 c = cv[r];
        if (act) {
          const double2 vi = vr[r];
          const double2 x = cmul(tk, pv[r]), y = cmul(al, vi);
          const double2 w = make_double2(x.x + y.x, x.y + y.y);
          wk[i] = w;
          const double2 t1 = cmul(vi, wkc), t2 = cmul(w, vkc);
          c.x -= t1.x + t2.x;
          c.y -= t1.y + t2.y;
        }
        cv[r] = c;
        if (i >= k + 3) nrm += c.x * c.x + c.y * c.y;
        if (i == k + 2) alph = c;
      }
    }
    // Y: this thread's rows of v_k^+ y (reflector k applied to y below)
    double2 sy = make_double2(0.0, 0.0);
    if (ys && act) {
#pragma unroll
      for (int r = 0; r < QS_HETRD_REG_RPT; ++r) {
        const int i = tid + r * bs;
        if (i > k && i < n) {
          const double2 t = cmulc(vk[i], yr[i]);
          sy.x += t.x;
          sy.y += t.y;
        }
      }
    }
    // the norm, the count of failed waits and v_k^+ y, over the workgroup:
    // one barrier (red is next written after the barrier ahead of the
    // owners' pass)
    {
      const double4_ q = {nrm, ok ? 0.0 : 1.0, sy.x, sy.y};
      const double t = wave_sum4(q.v);
      if ((lane & 15) == 0) red[4 * wv + (lane >> 4)] = t;
    }
    __syncthreads();
    // lane 8 q + w reads value q of wave w (one read each), the 8 lanes of
    // value q add them (DPP), every lane takes the four totals
    double xn2, nbad;
    {
      static_assert(QS_HETRD_REG_BLOCK == 512, "8 waves");
      const int q = lane >> 3, w = lane & 7;
      double x = q < 4 ? red[4 * w + q] : 0.0;
      x = sum8_lanes(x);
      xn2 = readlane_f64(x, 0);
      nbad = readlane_f64(x, 8);
      sy = make_double2(readlane_f64(x, 16), readlane_f64(x, 24));
    }
    if (nbad != 0.0) return;
    // y <- (I - tau_k v_k v_k^+)^+ y = y - conj(tau_k) (v_k^+ y) v_k, after
    // the owners' pass (nobody waits for it; v_k stays in LDS until step
    // k+1's first barrier), on the last step here
    auto apply_y = [&]() {
      if (!ys || !act) return;
      const double2 f = cmulc(tk, sy);
#pragma unroll
      for (int r = 0; r < QS_HETRD_REG_RPT; ++r) {
        const int i = tid + r * bs;
        if (i > k && i < n) {
          const double2 t = cmul(vk[i], f);
          yr[i].x -= t.x;
          yr[i].y -= t.y;
        }
      }
    };
    const double dk1 = cv[0].x;  // d[k+1]: row k+1 is thread 0's first
    if (k + 1 > n - 2) {
      if (lead && tid == 0) {
        d[k + 1] = dk1;
        A[(size_t)k * n + k + 1] = make_double2(ek, 0.0);
      }
      apply_y();
      break;
    }
    // the next reflector (LAPACK zlarfg on rows k+2..n-1 of column k+1): every
    // thread forms tau and beta from the norm and alpha, and writes its rows of
    // the reflector (leading 1, tail scaled) into vn
    double2 tn;
    {
      const double2 alpha = alph;
      double2 sc = make_double2(1.0, 0.0);
      if (xn2 == 0.0 && alpha.y == 0.0) {
        tn = make_double2(0.0, 0.0);
        beta = alpha.x;
      } else {
        const double nrm2 = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
        beta = alpha.x >= 0.0 ? -nrm2 : nrm2;
        // two reciprocals instead of four divisions (each a long dependent
        // sequence): tau = (beta - alpha) / beta, scale = 1 / (alpha - beta)
        const double rb = 1.0 / beta;
        tn = make_double2((beta - alpha.x) * rb, -alpha.y * rb);
        const double dr = alpha.x - beta, di = alpha.y;
        const double rdd = 1.0 / (dr * dr + di * di);
        sc = make_double2(dr * rdd, -di * rdd);
      }
#pragma unroll
      for (int r = 0; r < QS_HETRD_REG_RPT; ++r) {
        const int i = k + 1 + tid + r * bs;
        if (i < k + 2 || i >= n) continue;
        vn[i] = i == k + 2 ? make_double2(1.0, 0.0) : (tn.x != 0.0 || tn.y != 0.0) ? cmul(sc, cv[r]) : cv[r];
      }
    }
    const bool nact = tn.x != 0.0 || tn.y != 0.0;
    // the reflector is in LDS, and every thread's markers (issued at the end
    // of step k-1) have completed, before anything of step k+1 is stored (by
    // whichever thread)
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    owners(k + 2, act, nact, pbuf + (size_t)cw * n, part + (size_t)cw * G, colbuf + (size_t)cw * n);
    // off the critical path: reflector k on y, and into column k; T's entries
    apply_y();
    if (lead) {
      double2 *c0 = A + (size_t)k * n;
      for (int i = k + 2 + tid; i < n; i += bs) c0[i] = vk[i];
      if (tid == 0) {
        c0[k + 1] = make_double2(ek, 0.0);
        d[k + 1] = dk1;
        tau[k + 1] = tn;
        e[k + 1] = beta;
      }
    }
    // every workgroup has finished step k-1 (it has stored its partial of
    // step k): the marker back into this workgroup's words of copy k+2 mod 3
    // (its partial, its p entries, and column k+3 if it owns it), which step
    // k+1 writes. Issued last, they complete while step k+1 waits for its
    // data; the wait for them is the one ahead of step k+1's owners' pass.
    {
      double2 *pz = pbuf + (size_t)cz * n, *colz = colbuf + (size_t)cz * n;
      if (tid == 0) st_mark(&part[(size_t)cz * G + g]);
      if (tid < QS_HETRD_REG_NC && g + tid * G < n) st_mark(&pz[g + tid * G]);
      if (k + 3 < n && (k + 3) % G == g)
        for (int i = k + 3 + tid; i < n; i += bs) st_mark(&colz[i]);
    }
    // step k+1 overwrites vk (as its vn) and wk: the owners' pass has read
    // them before its own barrier when it forms products, the lead's copy of
    // the reflector needs one more
    if (!nact || lead) __syncthreads();
    {
      double2 *t = vk;
      vk = vn;
      vn = t;
    }
    tk = tn;
    ek = beta;
  }
  if (ys)
    for (int i = tid; i < n; i += bs) ys[i] = yr[i];
}

// In-place tridiagonal reduction of the n x n Hermitian matrix A (device,
// column-major, lower triangle read as zhetrd reads it) with Y <- Q^+ Y for nv
// vectors: fills d, e, tau (device). Returns QS_EUNSUPPORTED, leaving A and Y
// untouched, when the grid (one workgroup per QS_HETRD_REG_NC columns, at
// least nv, at most 256) does not fit the device, i.e. above 1536 rows on the
// MI355X; the caller then takes another route. Measured on the MI355X
// (eigenvalues and projections of 3 vectors for a random dense complex matrix,
// tridiagonal eigenproblem included): 0.83 against 2.0 ms with rocSOLVER's
// zhetrd at n = 120, 3.1 against 7.9 ms at 300, 8.0 against 20.3 ms at 800,
// 15.5 against 38.7 ms at 1410. Per column in the middle of the reduction
// (n = 1410, timestamps inside the kernel): 5.9 us, of which 2.4 waiting for
// the exchanged data (memory latency), 0.7 the next column and its norm, 0.5
// the reflector, 1.7 the owners' update and products (0.5 of it waiting at
// their barrier for the waves with more rows).
qs_status qs_hetrd_fused(qs_ctx *ctx, int n, double2 *A, double *d, double *e, double2 *tau, int nv, double2 *Y) {
  if (n < 1 || n > QS_HETRD_REG_MAX) return QS_EUNSUPPORTED;
  const size_t shm = (size_t)4 * n * sizeof(double2);
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
  long max_spins = (long)ctx->wait_spins;
  void *args[] = {&n, &A, &d, &e, &tau, &pbuf, &part, &colbuf, &nv, &Y, &err, &max_spins};
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
