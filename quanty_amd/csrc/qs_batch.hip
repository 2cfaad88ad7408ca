// Batched Lanczos: many independent recurrences (different operators, bases,
// start vectors and depths) advanced in lockstep by one set of launches per
// step. Each slot holds an assembled operator (sliced ELL or product layout),
// its two Lanczos vectors and its coefficient arrays; a product-layout slot
// keeps its vectors in the layout's row order (converted on the way in and
// out, so callers see canonical order); a device worklist of (slot, slice) and (slot, row
// chunk) entries drives the kernels, and per-slot partial sums are reduced in a
// fixed order, so a slot's coefficients are bit-identical whether it runs
// alone or among others. Step counters live on the device; a slot stops when
// it reaches its depth.
//
// Per step (recurrence of LanczosTriDiagonalizeCC, Core/Lanczos.cpp:2169-2264):
//   K1: h = H u_i, per-entry partial of Re<h|u_i>, w = s_i h - b_i s_{i-1} u_{i-1}
//   R1: a_{i+1} = s_i^2 * sum of the slot's partials
//   K2: u_{i+1} = w - a_{i+1} s_i u_i, partial |u_{i+1}|^2 (and y += z_i s_i u_i
//       when the slot rebuilds a Ritz vector)
//   R2: b_{i+1} = sqrt(sum), s_{i+1} = 1/b_{i+1}, step += 1, swap u_i/u_{i+1}
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include <map>
#include <tuple>

#include "qs_internal.h"
#include "qs_kernels.h"

#define QS_K2_CHUNK 4096  // rows per K2 work item

struct SlotDev {
  int64_t n;
  int64_t nslices;
  SellD A;               // the slot's assembled operator
  double2 *uc;
  double2 *up;
  double2 *y;            // Ritz-vector accumulator or null
  double2 *V;            // kept Lanczos vectors u_0..u_pad (n each) or null
  const double *ycoef;   // z_i, length depth
  double *a;             // depth + 1
  double *b;             // depth + 1
  double *s;             // depth + 1
  double sc;             // s_i of the current step i
  double cp;             // b_i s_{i-1} of the current step (0 at i = 0)
  int step;              // completed steps
  int depth;
  int active;
  int pad;               // capacity (maximum depth) of a, b, s, ycoef
};

struct WorkItem {
  int slot;
  int idx;  // group of QS_K1_GROUP slices (K1) or row chunk (K2)
};

#define QS_K1_GROUP 4  // slices per K1 work item: one per wavefront of a workgroup

struct qs_batch {
  qs_ctx *ctx = nullptr;
  int cap = 0;
  SlotDev *d_slots = nullptr;
  std::vector<SlotDev> h_slots;
  std::vector<int> used;
  std::vector<char *> mem;     // per-slot allocation (vectors + scalars)
  std::vector<size_t> mem_bytes;
  std::vector<double2 *> kept; // per-slot kept Lanczos vectors or null
  std::vector<size_t> kept_bytes;
  std::vector<const qs_hop *> hops;
  // worklists (rebuilt when the set of slots changes)
  bool dirty = true;
  WorkItem *d_w1 = nullptr, *d_w2 = nullptr;
  int64_t n1 = 0, n2 = 0, cap1 = 0, cap2 = 0;
  int64_t n1p = 0;  // K1 items of product-layout slots (the last n1p of the n1)
  int *d_r1 = nullptr, *d_r2 = nullptr;  // per slot: [begin, end) of its entries, 2 ints each
  int *d_off = nullptr;                    // per slot: first K1 item, first K2 item (worklist build)
  double *d_p1 = nullptr, *d_p2 = nullptr;  // partials per slice (K1) and per work item (K2)
  size_t lds1 = 0;                          // K1 dynamic LDS: the largest term table that is staged
  size_t ldsp = 0;                          // K1p: the largest term-value table that is staged
  bool any1 = false, anyp = false;          // sliced-ELL / product-layout slots present
  bool fast1 = true;                        // every sliced-ELL slot: one-word determinants, packed entries
  int64_t *d_n = nullptr;                   // work item counts of K1 and K2 (device, read by the kernels)
  // graphs of `steps` steps by (K1 variant and presence, K1 LDS, K1p presence, K1p LDS, steps);
  // valid while the worklist and partial buffers keep their addresses
  std::map<std::tuple<int, int, int, int, int>, hipGraphExec_t> graphs;
};

static void drop_graphs(qs_batch *B) {
  for (auto &g : B->graphs) (void)hipGraphExecDestroy(g.second);
  B->graphs.clear();
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(QS_BLOCK)
k_b_init(SlotDev *slots, int slot, const double2 *__restrict__ v0, double *__restrict__ part) {
  __shared__ double red[QS_BLOCK];
  SlotDev &S = slots[slot];
  double sum = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < S.n; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 x = v0[i];
    S.uc[i] = x;
    if (S.V) S.V[i] = x;
    S.up[i] = make_double2(0.0, 0.0);
    if (S.y) S.y[i] = make_double2(0.0, 0.0);
    sum += x.x * x.x + x.y * x.y;
  }
  const double t = block_sum_values(sum, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// A new slot's descriptor (passed by value, so no host copy is staged) and
// its zeroed coefficient arrays.
__global__ void k_b_slot_set(SlotDev *slots, int slot, SlotDev S, double *__restrict__ scal, int nscal) {
  for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < nscal; i += (int)(gridDim.x * blockDim.x))
    scal[i] = 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) slots[slot] = S;
}

__global__ void __launch_bounds__(QS_BLOCK)
k_b_init_finish(SlotDev *slots, int slot, const double *__restrict__ part, int np) {
  __shared__ double red[QS_BLOCK];
  const double nrm2 = block_sum_fixed(part, np, red);
  if (threadIdx.x == 0) {
    SlotDev &S = slots[slot];
    const double b0 = sqrt(nrm2);
    S.a[0] = 0.0;
    S.b[0] = b0;
    S.s[0] = b0 > 0.0 ? 1.0 / b0 : 0.0;
    S.sc = S.s[0];
    S.cp = 0.0;
    S.step = 0;
  }
}

// K1 over the work items of all active sliced-ELL slots (k_b1p takes the
// product-layout ones). A workgroup takes one item (4 consecutive slices of
// one slot, one per wavefront) per iteration and keeps the slot's term tables
// in LDS while consecutive items stay in that slot; tables larger than the
// dynamic LDS are read from global memory. The items
// of XCD x (workgroups b with b % 8 == x) are a contiguous range, so a
// workgroup meets few slot changes.
template <int NW, bool PACKED>
__device__ __forceinline__ void b1_row(const SellD &A, int64_t sl, int lane, bool lds, const char *smem,
                                       const double2 *uc, double *hr, double *hi) {
  if (lds)
    sell_row<1, NW, PACKED>(A, sl, lane, reinterpret_cast<const double2 *>(smem),
                            reinterpret_cast<const uint64_t *>(smem + (size_t)A.nterms * 16), uc, 1, 0, hr, hi);
  else
    sell_row<1, NW, PACKED>(A, sl, lane, A.val, A.par, uc, 1, 0, hr, hi);
}

// FAST: every active sliced-ELL slot has one-word determinants and packed
// entries (the common case), which keeps the kernel within 64 VGPRs (8 waves
// per SIMD).
template <bool FAST>
__global__ void __launch_bounds__(QS_BLOCK, FAST ? 8 : 1)
k_b1(const SlotDev *__restrict__ slots, const WorkItem *__restrict__ work, const int64_t *__restrict__ counts,
     double *__restrict__ part, int lds_bytes) {
  extern __shared__ char smem[];
  const int64_t nwork = counts[2];  // the sliced-ELL slots' items come first
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  constexpr int NX = 8;
  const int64_t nblk = gridDim.x;
  int64_t i0, i1, di;
  if (nblk < NX) {
    i0 = blockIdx.x;
    i1 = nwork;
    di = nblk;
  } else {
    const int xcd = blockIdx.x % NX;
    i0 = nwork * xcd / NX + blockIdx.x / NX;
    i1 = nwork * (xcd + 1) / NX;
    di = (nblk - xcd + NX - 1) / NX;
  }
  int staged = -1;
  for (int64_t i = i0; i < i1; i += di) {
    const WorkItem it = work[i];
    const SlotDev &S = slots[it.slot];
    if (!S.active || S.step >= S.depth) continue;  // uniform across the workgroup
    if (S.A.kind != 0) continue;  // k_b1p's item; uniform across the workgroup
    // the item's operator, vectors and scalars as locals
    const SellD A = S.A;
    double2 *const uc = S.uc, *const up = S.up;
    const int64_t n = S.n, nsl = S.nslices;
    const double sc = S.sc, cp = S.cp;
    const bool lds = term_table_bytes(A.nterms, A.nw) <= (size_t)lds_bytes;
    if (lds && staged != it.slot) {
      __syncthreads();  // the previous table is no longer read
      if (FAST || A.nw == 1) stage_term_tables<1>(A, smem);
      else stage_term_tables<2>(A, smem);
      __syncthreads();
      staged = it.slot;
    }
    const int64_t sl = (int64_t)it.idx * QS_K1_GROUP + wv;
    double dot = 0.0;
    if (sl < nsl) {  // uniform across the wavefront
      double hr[1], hi[1];
      if (FAST) {
        b1_row<1, true>(A, sl, lane, lds, smem, uc, hr, hi);
      } else if (A.nw == 1) {
        if (A.packed) b1_row<1, true>(A, sl, lane, lds, smem, uc, hr, hi);
        else b1_row<1, false>(A, sl, lane, lds, smem, uc, hr, hi);
      } else {
        if (A.packed) b1_row<2, true>(A, sl, lane, lds, smem, uc, hr, hi);
        else b1_row<2, false>(A, sl, lane, lds, smem, uc, hr, hi);
      }
      const int64_t r = sl * 64 + lane;
      if (r < n) {
        const double2 u = uc[r];
        dot = hr[0] * u.x + hi[0] * u.y;
        const double2 p = up[r];
        up[r] = make_double2(sc * hr[0] - cp * p.x, sc * hi[0] - cp * p.y);
      }
    }
    dot = wave_sum(dot);  // fixed butterfly: same order every call
    if (lane == 0) part[i * QS_K1_GROUP + wv] = dot;
  }
}

// K1 over the work items of the product-layout slots, whose vectors are in
// the layout's row order: the slot's term values in LDS when they fit in
// lds_bytes, then a staging area of QS_PROD_STAGE vector entries per
// wavefront for the rows of a slice's group.
__global__ void __launch_bounds__(QS_BLOCK)
k_b1p(const SlotDev *__restrict__ slots, const WorkItem *__restrict__ work, const int64_t *__restrict__ counts,
      double *__restrict__ part, int lds_bytes) {
  extern __shared__ char smem[];
  const int64_t w0 = counts[2], nwork = counts[0] - w0;  // the product slots' items come last
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  constexpr int NX = 8;
  const int64_t nblk = gridDim.x;
  int64_t i0, i1, di;
  if (nblk < NX) {
    i0 = w0 + blockIdx.x;
    i1 = w0 + nwork;
    di = nblk;
  } else {
    const int xcd = blockIdx.x % NX;
    i0 = w0 + nwork * xcd / NX + blockIdx.x / NX;
    i1 = w0 + nwork * (xcd + 1) / NX;
    di = (nblk - xcd + NX - 1) / NX;
  }
  double2 *xs = reinterpret_cast<double2 *>(smem + lds_bytes) + (size_t)wv * QS_PROD_STAGE;
  int staged = -1;
  for (int64_t i = i0; i < i1; i += di) {
    const WorkItem it = work[i];
    const SlotDev &S = slots[it.slot];
    if (!S.active || S.step >= S.depth) continue;  // uniform across the workgroup
    if (S.A.kind != 1) continue;  // k_b1's item
    // the item's operator, vectors and scalars as locals
    const SellD A = S.A;
    double2 *const uc = S.uc, *const up = S.up;
    const int64_t nsl = S.nslices;
    const double sc = S.sc, cp = S.cp;
    const bool lds = term_value_bytes(A.nterms) <= (size_t)lds_bytes;
    if (lds && staged != it.slot) {
      __syncthreads();  // the previous values are no longer read
      stage_term_values(A, smem);
      __syncthreads();
      staged = it.slot;
    }
    const int64_t sl = (int64_t)it.idx * QS_K1_GROUP + wv;
    double dot = 0.0;
    if (sl < nsl) {  // uniform across the wavefront
      double hr[1], hi[1];
      // separate calls, so the term values are read as LDS or as global memory
      // (one pointer to either would make every read a flat load)
      const int64_t r = lds ? prod_row<1>(A, sl, lane, reinterpret_cast<const double2 *>(smem), uc, 1, 0, hr, hi, xs)
                            : prod_row<1>(A, sl, lane, A.val, uc, 1, 0, hr, hi, xs);
      if (r >= 0) {
        const double2 u = uc[r];
        dot = hr[0] * u.x + hi[0] * u.y;
        const double2 p = up[r];
        up[r] = make_double2(sc * hr[0] - cp * p.x, sc * hi[0] - cp * p.y);
      }
    }
    dot = wave_sum(dot);  // fixed butterfly: same order every call
    if (lane == 0) part[i * QS_K1_GROUP + wv] = dot;
  }
}

// Sum of a slot's work-item partials part[b..e) by one 256-thread block, in
// an order fixed by e - b alone (thread t takes b+t, b+t+256, ... into four
// interleaved accumulators), so a slot's result does not depend on the other
// slots of the batch.
__device__ __forceinline__ double slot_sum(const double *__restrict__ part, int b, int e, double *red) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int k = b + threadIdx.x;
  const int st = blockDim.x;
  for (; k + 3 * st < e; k += 4 * st) {
    s0 += part[k];
    s1 += part[k + st];
    s2 += part[k + 2 * st];
    s3 += part[k + 3 * st];
  }
  for (; k < e; k += st) s0 += part[k];
  return block_sum_values((s0 + s1) + (s2 + s3), red);
}

// one block per slot: a_{i+1} = s_i^2 Re<H u_i|u_i>. A large slot has one
// partial per slice (34k at 2.2 M rows): 1024 threads keep the sum short.
#define QS_R1_BLOCK 1024
__global__ void __launch_bounds__(QS_R1_BLOCK)
k_b_r1(SlotDev *slots, int nslots, const int *__restrict__ range, const double *__restrict__ part) {
  __shared__ double red[QS_R1_BLOCK / 64];
  const int slot = blockIdx.x;
  if (slot >= nslots) return;
  SlotDev &S = slots[slot];
  if (!S.active || S.step >= S.depth) return;  // uniform across the block
  const double sum = slot_sum(part, range[2 * slot], range[2 * slot + 1], red);
  if (threadIdx.x == 0) {
    const double sc = S.s[S.step];
    S.a[S.step + 1] = sc * sc * sum;
  }
}

__global__ void __launch_bounds__(QS_BLOCK)
k_b2(const SlotDev *__restrict__ slots, const WorkItem *__restrict__ work, const int64_t *__restrict__ counts,
     double *__restrict__ part) {
  __shared__ double red[QS_BLOCK];
  const int64_t nwork = counts[1];
  for (int64_t w = blockIdx.x; w < nwork; w += gridDim.x) {
    const WorkItem it = work[w];
    const SlotDev &S = slots[it.slot];
    if (!S.active || S.step >= S.depth) continue;  // uniform across the block
    const int i = S.step;
    const double sc = S.s[i];
    const double f = S.a[i + 1] * sc;
    const double zc = S.y ? S.ycoef[i] * sc : 0.0;
    const int64_t r0 = (int64_t)it.idx * QS_K2_CHUNK;
    const int64_t r1 = min(S.n, r0 + QS_K2_CHUNK);
    double nrm = 0.0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
      const double2 wv = S.up[r];
      const double2 u = S.uc[r];
      if (S.y) {
        const double2 yv = S.y[r];
        S.y[r] = make_double2(fma(zc, u.x, yv.x), fma(zc, u.y, yv.y));
      }
      const double2 un = make_double2(wv.x - f * u.x, wv.y - f * u.y);
      S.up[r] = un;
      if (S.V) S.V[(size_t)(i + 1) * S.n + r] = un;
      nrm += un.x * un.x + un.y * un.y;
    }
    const double t = block_sum_values(nrm, red);
    if (threadIdx.x == 0) part[w] = t;
  }
}

__global__ void __launch_bounds__(QS_BLOCK)
k_b_r2(SlotDev *slots, int nslots, const int *__restrict__ range, const double *__restrict__ part) {
  __shared__ double red[QS_BLOCK / 64];
  const int slot = blockIdx.x;
  if (slot >= nslots) return;
  SlotDev &S = slots[slot];
  if (!S.active || S.step >= S.depth) return;  // uniform across the block
  const double sum = slot_sum(part, range[2 * slot], range[2 * slot + 1], red);
  if (threadIdx.x == 0) {
    const int i = S.step + 1;
    const double bn = sqrt(sum);
    S.b[i] = bn;
    S.s[i] = bn > 0.0 ? 1.0 / bn : 0.0;
    S.sc = S.s[i];
    S.cp = bn * S.s[i - 1];
    double2 *t = S.uc;
    S.uc = S.up;
    S.up = t;
    S.step = i;
  }
}

// y = sum_{j<m} z_j s_j u_j over the kept vectors, j ascending per row (the
// order of the second-pass accumulation in k_b2); row r goes to y[inv[r]]
// when the slot keeps its vectors in a product layout's row order
__global__ void __launch_bounds__(QS_BLOCK)
k_b_ritz(const SlotDev *__restrict__ slots, int slot, int m, const double *__restrict__ z,
         const int32_t *__restrict__ inv, double2 *__restrict__ y) {
  const SlotDev &S = slots[slot];
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < S.n; r += (int64_t)gridDim.x * blockDim.x) {
    double2 acc = make_double2(0.0, 0.0);
    for (int j = 0; j < m; ++j) {
      const double c = z[j] * S.s[j];
      const double2 u = S.V[(size_t)j * S.n + r];
      acc = make_double2(fma(c, u.x, acc.x), fma(c, u.y, acc.y));
    }
    y[inv ? (int64_t)inv[r] : r] = acc;
  }
}

// Worklists from the slot table, on the device, so a change in the set of
// slots needs no host upload and no wait for the steps in flight (the stream
// orders the build after them). A slot's K1 items are its groups of
// QS_K1_GROUP slices, its K2 items its row chunks, slots in index order.
// Scan: one block; thread t of each tile of blockDim.x slots counts slot
// base + t, and an inclusive scan in LDS gives the offsets.
#define QS_WL_BLOCK 1024
__global__ void __launch_bounds__(QS_WL_BLOCK)
k_b_worklist_scan(const SlotDev *__restrict__ slots, int cap, int *__restrict__ off, int *__restrict__ r1,
                  int *__restrict__ r2, int64_t *__restrict__ counts) {
  __shared__ int s1[QS_WL_BLOCK], s2[QS_WL_BLOCK];
  int base1 = 0, base2 = 0;
  // K1 items: the sliced-ELL slots' first (pass 0), then the product-layout
  // slots' (pass 1), so k_b1 and k_b1p each walk a range of their own; K2
  // items in pass 0
  for (int pass = 0; pass < 2; ++pass) {
    if (pass == 1 && threadIdx.x == 0) counts[2] = base1;
    for (int t0 = 0; t0 < cap; t0 += blockDim.x) {
      const int s = t0 + (int)threadIdx.x;
      const bool act = s < cap && slots[s].active;
      const bool mine = s < cap && (act ? (slots[s].A.kind == 1) == (pass == 1) : pass == 0);
      int c1 = 0, c2 = 0;
      if (act && mine) c1 = (int)((slots[s].nslices + QS_K1_GROUP - 1) / QS_K1_GROUP);
      if (act && pass == 0) c2 = (int)((slots[s].n + QS_K2_CHUNK - 1) / QS_K2_CHUNK);
      s1[threadIdx.x] = c1;
      s2[threadIdx.x] = c2;
      __syncthreads();
      for (int o = 1; o < (int)blockDim.x; o <<= 1) {
        const int a1 = threadIdx.x >= (unsigned)o ? s1[threadIdx.x - o] : 0;
        const int a2 = threadIdx.x >= (unsigned)o ? s2[threadIdx.x - o] : 0;
        __syncthreads();
        s1[threadIdx.x] += a1;
        s2[threadIdx.x] += a2;
        __syncthreads();
      }
      if (mine) {
        const int e1 = base1 + s1[threadIdx.x];
        off[2 * s] = e1 - c1;
        // K1 partials are kept per slice position of an item
        r1[2 * s] = (e1 - c1) * QS_K1_GROUP;
        r1[2 * s + 1] = e1 * QS_K1_GROUP;
      }
      if (s < cap && pass == 0) {
        const int e2 = base2 + s2[threadIdx.x];
        off[2 * s + 1] = e2 - c2;
        r2[2 * s] = e2 - c2;
        r2[2 * s + 1] = e2;
      }
      base1 += s1[blockDim.x - 1];
      base2 += s2[blockDim.x - 1];
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) {
    counts[0] = base1;
    counts[1] = base2;
  }
}

__global__ void __launch_bounds__(QS_BLOCK)
k_b_worklist_fill(const SlotDev *__restrict__ slots, const int *__restrict__ off, WorkItem *__restrict__ w1,
                  WorkItem *__restrict__ w2) {
  const int s = blockIdx.x;
  if (!slots[s].active) return;
  const int c1 = (int)((slots[s].nslices + QS_K1_GROUP - 1) / QS_K1_GROUP);
  const int c2 = (int)((slots[s].n + QS_K2_CHUNK - 1) / QS_K2_CHUNK);
  const int o1 = off[2 * s], o2 = off[2 * s + 1];
  for (int k = threadIdx.x; k < c1; k += blockDim.x) w1[o1 + k] = WorkItem{s, k};
  for (int k = threadIdx.x; k < c2; k += blockDim.x) w2[o2 + k] = WorkItem{s, k};
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
// Slot memory comes from the library's device memory cache, which keeps
// freed blocks for reuse across slots, batches and servers.
static qs_status pool_get(qs_batch *B, size_t need, void **p, size_t *bytes) {
  QS_TRY(qs_malloc(B->ctx, p, need));
  *bytes = need;
  return QS_OK;
}

static void pool_put(qs_batch *B, void *p, size_t bytes) {
  (void)B;
  (void)bytes;
  qs_free(p);
}

static qs_status rebuild(qs_batch *B) {
  qs_ctx *ctx = B->ctx;
  // item counts from the host's copy of the slot table (the device builds
  // the same lists from its own copy)
  int64_t n1 = 0, n2 = 0, n1p = 0;
  size_t lds = 0, ldsp = 0;
  bool fast = true, any1 = false, anyp = false;
  for (int s = 0; s < B->cap; ++s)
    if (B->used[s]) {
      const SlotDev &S = B->h_slots[s];
      n1 += (S.nslices + QS_K1_GROUP - 1) / QS_K1_GROUP;
      n2 += (S.n + QS_K2_CHUNK - 1) / QS_K2_CHUNK;
      const SellD &A = S.A;
      if (A.kind == 1) {
        n1p += (S.nslices + QS_K1_GROUP - 1) / QS_K1_GROUP;
        const size_t tb = (term_value_bytes(A.nterms) + 15) / 16 * 16;
        if (tb <= (size_t)ctx->term_lds) ldsp = std::max(ldsp, tb);
        anyp = true;
      } else {
        const size_t tb = term_table_bytes(A.nterms, A.nw);
        if (tb <= (size_t)ctx->term_lds) lds = std::max(lds, tb);
        fast = fast && A.nw == 1 && A.packed;
        any1 = true;
      }
    }
  if (n1 >= ((int64_t)1 << 31) / QS_K1_GROUP || n2 >= ((int64_t)1 << 31))
    return qs_fail(ctx, QS_EINVAL, "the batch's work lists would exceed 2^31 items");
  B->lds1 = lds;
  B->ldsp = ldsp;
  B->n1p = n1p;
  B->any1 = any1;
  B->anyp = anyp;
  B->fast1 = fast;
  auto grow = [&](WorkItem **p, int64_t *cap, int64_t need) -> qs_status {
    if (need > *cap) {
      qs_free(*p);
      *p = nullptr;
      int64_t c = std::max<int64_t>(need, 2 * *cap);
      QS_TRY(qs_malloc(ctx, (void **)p, (size_t)c * sizeof(WorkItem)));
      *cap = c;
    }
    return QS_OK;
  };
  const int64_t c1 = B->cap1, c2 = B->cap2;
  const void *old_ptrs[4] = {B->d_w1, B->d_w2, B->d_p1, B->d_p2};
  QS_TRY(grow(&B->d_w1, &B->cap1, std::max<int64_t>(1, n1)));
  QS_TRY(grow(&B->d_w2, &B->cap2, std::max<int64_t>(1, n2)));
  if (B->cap1 != c1 || !B->d_p1) {
    qs_free(B->d_p1);
    QS_TRY(qs_malloc(ctx, (void **)&B->d_p1, (size_t)B->cap1 * QS_K1_GROUP * 8));
  }
  if (B->cap2 != c2 || !B->d_p2) {
    qs_free(B->d_p2);
    QS_TRY(qs_malloc(ctx, (void **)&B->d_p2, (size_t)B->cap2 * 8));
  }
  hipLaunchKernelGGL(k_b_worklist_scan, dim3(1), dim3(QS_WL_BLOCK), 0, ctx->stream, B->d_slots, B->cap, B->d_off,
                     B->d_r1, B->d_r2, B->d_n);
  hipLaunchKernelGGL(k_b_worklist_fill, dim3(B->cap), dim3(QS_BLOCK), 0, ctx->stream, B->d_slots, B->d_off,
                     B->d_w1, B->d_w2);
  QS_HIP(ctx, hipGetLastError());
  if (old_ptrs[0] != B->d_w1 || old_ptrs[1] != B->d_w2 || old_ptrs[2] != B->d_p1 || old_ptrs[3] != B->d_p2)
    drop_graphs(B);
  B->n1 = n1;
  B->n2 = n2;
  B->dirty = false;
  return QS_OK;
}

extern "C" qs_status qs_batch_create(qs_ctx *ctx, int max_slots, qs_batch **out) {
  if (!ctx || !out || max_slots < 1) return QS_EINVAL;
  *out = nullptr;
  QS_HIP(ctx, hipSetDevice(ctx->device));
  qs_batch *B = new qs_batch();
  B->ctx = ctx;
  B->cap = max_slots;
  B->h_slots.assign(max_slots, SlotDev{});
  B->used.assign(max_slots, 0);
  B->mem.assign(max_slots, nullptr);
  B->mem_bytes.assign(max_slots, 0);
  B->kept.assign(max_slots, nullptr);
  B->kept_bytes.assign(max_slots, 0);
  B->hops.assign(max_slots, nullptr);
  qs_status st = qs_malloc(ctx, (void **)&B->d_slots, (size_t)max_slots * sizeof(SlotDev));
  if (st == QS_OK) st = qs_malloc(ctx, (void **)&B->d_r1, (size_t)max_slots * 8);
  if (st == QS_OK) st = qs_malloc(ctx, (void **)&B->d_r2, (size_t)max_slots * 8);
  if (st == QS_OK) st = qs_malloc(ctx, (void **)&B->d_n, 4 * sizeof(int64_t));
  if (st == QS_OK) st = qs_malloc(ctx, (void **)&B->d_off, (size_t)max_slots * 8);
  if (st == QS_OK) {
    hipError_t e = hipMemsetAsync(B->d_slots, 0, (size_t)max_slots * sizeof(SlotDev), ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) st = qs_hip_fail(ctx, e, "qs_batch_create", __FILE__, __LINE__);
  }
  if (st != QS_OK) {
    qs_batch_destroy(B);
    return st;
  }
  *out = B;
  return QS_OK;
}

extern "C" qs_status qs_batch_add(qs_batch *B, const qs_hop *h, const double *v0, int depth, int max_depth,
                                  const double *ycoef, int *slot_out) {
  return qs_batch_add_ex(B, h, v0, depth, max_depth, 0, ycoef, slot_out);
}

extern "C" qs_status qs_batch_add_ex(qs_batch *B, const qs_hop *h, const double *v0, int depth, int max_depth,
                                     int flags, const double *ycoef, int *slot_out) {
  if (!B || !h || !v0 || depth < 1 || !slot_out) return QS_EINVAL;
  if (max_depth < depth) max_depth = depth;
  qs_ctx *ctx = B->ctx;
  if (!h->assembled || h->rows != h->cols)
    return qs_fail(ctx, QS_EINVAL, "qs_batch_add needs an assembled operator with rows == cols");
  int slot = -1;
  for (int s = 0; s < B->cap; ++s)
    if (!B->used[s]) {
      slot = s;
      break;
    }
  if (slot < 0) return qs_fail(ctx, QS_EINVAL, "batch is full (%d slots)", B->cap);
  const int64_t n = h->rows->n;
  const size_t vb = (size_t)std::max<int64_t>(n, 1) * 16;
  const bool want_y = ycoef != nullptr;
  const size_t sc = (size_t)(max_depth + 1) * 8;
  if (flags & ~QS_BATCH_KEEP_BASIS) return qs_fail(ctx, QS_EINVAL, "unknown flags 0x%x", flags);
  const bool keep = (flags & QS_BATCH_KEEP_BASIS) != 0;
  char *m = nullptr;
  size_t mbytes = 0;
  QS_TRY(pool_get(B, (want_y ? 3 : 2) * vb + 4 * sc + 256, (void **)&m, &mbytes));
  double2 *V = nullptr;
  size_t vbytes = 0;
  if (keep) {
    qs_status sv = pool_get(B, (size_t)(max_depth + 1) * vb, (void **)&V, &vbytes);
    if (sv != QS_OK) {
      pool_put(B, m, mbytes);
      return sv;
    }
  }
  SlotDev S{};
  S.V = V;
  S.n = n;
  S.nslices = h->nslices;
  S.A = h->view;
  S.uc = (double2 *)m;
  S.up = (double2 *)(m + vb);
  S.y = want_y ? (double2 *)(m + 2 * vb) : nullptr;
  double *scal = (double *)(m + (want_y ? 3 : 2) * vb);
  S.a = scal;
  S.b = scal + (max_depth + 1);
  S.s = scal + 2 * (max_depth + 1);
  double *yc = scal + 3 * (max_depth + 1);
  S.ycoef = want_y ? yc : nullptr;
  S.step = 0;
  S.depth = depth;
  S.active = 1;
  S.pad = max_depth;  // capacity of the coefficient arrays
  hipError_t e = hipSuccess;
  qs_status st = QS_OK;
  if (want_y) st = qs_h2d(ctx, yc, ycoef, (size_t)depth * 8);
  if (st == QS_OK) {
    const int nscal = 3 * (max_depth + 1);
    hipLaunchKernelGGL(k_b_slot_set, dim3((unsigned)std::min(64, (nscal + QS_BLOCK - 1) / QS_BLOCK)), dim3(QS_BLOCK),
                       0, ctx->stream, B->d_slots, slot, S, scal, nscal);
    e = hipGetLastError();
  }
  const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(256, (n + QS_BLOCK - 1) / QS_BLOCK));
  double *part = nullptr;
  void *scr = nullptr;
  if (st == QS_OK) st = qs_scratch(ctx, (size_t)nb * 8, &scr);
  part = (double *)scr;
  // a product-layout slot starts from v0 in the layout's row order (staged in
  // its second vector, which k_b_init reads before clearing)
  const double2 *v0s = (const double2 *)v0;
  if (st == QS_OK && e == hipSuccess && h->view.kind == 1) {
    st = qs_prod_to_rows(ctx, h, 1, (const double2 *)v0, n, S.up, n);
    v0s = S.up;
  }
  if (st == QS_OK && e == hipSuccess) {
    hipLaunchKernelGGL(k_b_init, dim3(nb), dim3(QS_BLOCK), 0, ctx->stream, B->d_slots, slot, v0s, part);
    hipLaunchKernelGGL(k_b_init_finish, dim3(1), dim3(QS_BLOCK), 0, ctx->stream, B->d_slots, slot, part, nb);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (st != QS_OK || e != hipSuccess) {
    (void)hipStreamSynchronize(ctx->stream);
    pool_put(B, m, mbytes);
    pool_put(B, V, vbytes);
    return st != QS_OK ? st : qs_hip_fail(ctx, e, "qs_batch_add", __FILE__, __LINE__);
  }
  B->h_slots[slot] = S;
  B->used[slot] = 1;
  B->mem[slot] = m;
  B->mem_bytes[slot] = mbytes;
  B->kept[slot] = V;
  B->kept_bytes[slot] = vbytes;
  B->hops[slot] = h;
  B->dirty = true;
  *slot_out = slot;
  return QS_OK;
}

// K1p's dynamic LDS: the term values, then the staging area
static int kp_shm(const qs_batch *B) {
  return (int)B->ldsp + (QS_BLOCK / 64) * QS_PROD_STAGE * (int)sizeof(double2);
}

// `steps` steps as a graph of kernel nodes in a chain (the same launches as
// qs_batch_run's kernel-by-kernel path, with grids that fit any worklist).
static qs_status build_graph(qs_batch *B, int steps, int g1, int gp, int g2, hipGraphExec_t *out) {
  qs_ctx *ctx = B->ctx;
  hipGraph_t g = nullptr;
  QS_HIP(ctx, hipGraphCreate(&g, 0));
  const SlotDev *cs = B->d_slots;
  SlotDev *ms = B->d_slots;
  const WorkItem *w1 = B->d_w1, *w2 = B->d_w2;
  const int64_t *cn = B->d_n;
  double *p1 = B->d_p1, *p2 = B->d_p2;
  const double *cp1 = B->d_p1, *cp2 = B->d_p2;
  const int *r1 = B->d_r1, *r2 = B->d_r2;
  int cap = B->cap, lds = (int)B->lds1, ldsp = (int)B->ldsp;
  void *a1[] = {&cs, &w1, &cn, &p1, &lds};
  void *ap[] = {&cs, &w1, &cn, &p1, &ldsp};
  void *ar1[] = {&ms, &cap, &r1, &cp1};
  void *a2[] = {&cs, &w2, &cn, &p2};
  void *ar2[] = {&ms, &cap, &r2, &cp2};
  void *k1 = B->fast1 ? reinterpret_cast<void *>(&k_b1<true>) : reinterpret_cast<void *>(&k_b1<false>);
  struct Node {
    void *f;
    int grid;
    int block;
    int shm;
    void **args;
  };
  // per step: K1 and K1p side by side (both after the previous step, R1 after
  // both, so the tail of one overlaps the other), then R1, K2, R2 in a chain
  std::vector<Node> k1s, chain;
  if (B->any1) k1s.push_back({k1, g1, QS_BLOCK, lds, a1});
  if (B->anyp) k1s.push_back({reinterpret_cast<void *>(&k_b1p), gp, QS_BLOCK, kp_shm(B), ap});
  chain.push_back({reinterpret_cast<void *>(&k_b_r1), B->cap, QS_R1_BLOCK, 0, ar1});
  chain.push_back({reinterpret_cast<void *>(&k_b2), g2, QS_BLOCK, 0, a2});
  chain.push_back({reinterpret_cast<void *>(&k_b_r2), B->cap, QS_BLOCK, 0, ar2});
  auto add = [&](const Node &nd, const hipGraphNode_t *deps, size_t ndeps, hipGraphNode_t *node) {
    hipKernelNodeParams p{};
    p.func = nd.f;
    p.gridDim = dim3((unsigned)nd.grid);
    p.blockDim = dim3((unsigned)nd.block);
    p.sharedMemBytes = (unsigned)nd.shm;
    p.kernelParams = nd.args;
    p.extra = nullptr;
    return hipGraphAddKernelNode(node, g, deps, ndeps, &p);
  };
  hipGraphNode_t prev = nullptr;
  hipError_t e = hipSuccess;
  for (int k = 0; k < steps && e == hipSuccess; ++k) {
    hipGraphNode_t k1n[2];
    size_t nk = 0;
    for (const Node &nd : k1s) {
      e = add(nd, prev ? &prev : nullptr, prev ? 1 : 0, &k1n[nk]);
      if (e != hipSuccess) break;
      ++nk;
    }
    for (size_t c = 0; c < chain.size() && e == hipSuccess; ++c) {
      hipGraphNode_t node = nullptr;
      if (c == 0) e = add(chain[c], nk ? k1n : (prev ? &prev : nullptr), nk ? nk : (prev ? 1 : 0), &node);
      else e = add(chain[c], &prev, 1, &node);
      prev = node;
    }
  }
  if (e == hipSuccess) e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return qs_hip_fail(ctx, e, "qs_batch_run (graph)", __FILE__, __LINE__);
  return QS_OK;
}

extern "C" qs_status qs_batch_run(qs_batch *B, int steps) {
  if (!B || steps < 0) return QS_EINVAL;
  qs_ctx *ctx = B->ctx;
  QS_HIP(ctx, hipSetDevice(ctx->device));
  if (B->dirty) QS_TRY(rebuild(B));
  if (B->n1 == 0) return QS_OK;
  const int lds1 = (int)B->lds1, ldsp = (int)B->ldsp, shmp = kp_shm(B);
  const void *k1 = B->fast1 ? (const void *)k_b1<true> : (const void *)k_b1<false>;
  const int64_t res1 = qs_resident_blocks(ctx, k1, (size_t)lds1);
  const int64_t resp = qs_resident_blocks(ctx, (const void *)k_b1p, (size_t)shmp);
  const int64_t res2 = (int64_t)ctx->num_cu * 4;
  // the kernels read the work item counts from the device, so any grid
  // covers the worklist; a graph uses the resident grids, which fit every
  // later worklist
  auto enqueue = [&](int g1, int gp, int g2, int n) {
    for (int k = 0; k < n; ++k) {
      if (B->any1) {
        if (B->fast1)
          hipLaunchKernelGGL(k_b1<true>, dim3(g1), dim3(QS_BLOCK), (size_t)lds1, ctx->stream, B->d_slots, B->d_w1,
                             B->d_n, B->d_p1, lds1);
        else
          hipLaunchKernelGGL(k_b1<false>, dim3(g1), dim3(QS_BLOCK), (size_t)lds1, ctx->stream, B->d_slots, B->d_w1,
                             B->d_n, B->d_p1, lds1);
      }
      if (B->anyp)
        hipLaunchKernelGGL(k_b1p, dim3(gp), dim3(QS_BLOCK), (size_t)shmp, ctx->stream, B->d_slots, B->d_w1, B->d_n,
                           B->d_p1, ldsp);
      hipLaunchKernelGGL(k_b_r1, dim3(B->cap), dim3(QS_R1_BLOCK), 0, ctx->stream, B->d_slots, B->cap, B->d_r1,
                         B->d_p1);
      hipLaunchKernelGGL(k_b2, dim3(g2), dim3(QS_BLOCK), 0, ctx->stream, B->d_slots, B->d_w2, B->d_n, B->d_p2);
      hipLaunchKernelGGL(k_b_r2, dim3(B->cap), dim3(QS_BLOCK), 0, ctx->stream, B->d_slots, B->cap, B->d_r2, B->d_p2);
    }
  };
  if (ctx->batch_graph && steps > 0) {
    const auto key = std::make_tuple((B->any1 ? 1 : 0) + (B->fast1 ? 2 : 0), lds1, B->anyp ? 1 : 0, ldsp, steps);
    auto it = B->graphs.find(key);
    if (it == B->graphs.end()) {
      // The graph is assembled node by node rather than captured from the
      // stream: other host threads may synchronize the device meanwhile,
      // which would invalidate a capture.
      hipGraphExec_t ex = nullptr;
      QS_TRY(build_graph(B, steps, (int)std::max<int64_t>(1, res1), (int)std::max<int64_t>(1, resp),
                         (int)std::max<int64_t>(1, res2), &ex));
      it = B->graphs.emplace(key, ex).first;
    }
    QS_HIP(ctx, hipGraphLaunch(it->second, ctx->stream));
  } else {
    enqueue((int)std::max<int64_t>(1, std::min(B->n1 - B->n1p, res1)), (int)std::max<int64_t>(1, std::min(B->n1p, resp)),
            (int)std::max<int64_t>(1, std::min(B->n2, res2)), steps);
    QS_HIP(ctx, hipGetLastError());
  }
  if (!ctx->async) QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QS_OK;
}

__global__ void k_b_set_depth(SlotDev *slots, int slot, int depth) { slots[slot].depth = depth; }

extern "C" qs_status qs_batch_set_depth(qs_batch *B, int slot, int depth) {
  if (!B || slot < 0 || slot >= B->cap || !B->used[slot]) return QS_EINVAL;
  qs_ctx *ctx = B->ctx;
  SlotDev &H = B->h_slots[slot];
  if (depth < 1 || depth > H.pad) return qs_fail(ctx, QS_EINVAL, "depth %d outside 1..%d", depth, H.pad);
  if (H.y) return qs_fail(ctx, QS_EINVAL, "the depth of a vector-building slot is fixed");
  QS_HIP(ctx, hipSetDevice(ctx->device));
  hipLaunchKernelGGL(k_b_set_depth, dim3(1), dim3(1), 0, ctx->stream, B->d_slots, slot, depth);
  QS_HIP(ctx, hipGetLastError());
  QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  H.depth = depth;
  return QS_OK;
}

extern "C" qs_status qs_batch_poll(qs_batch *B, int slot, int *step, double *a, double *b, double *y) {
  if (!B || slot < 0 || slot >= B->cap || !B->used[slot]) return QS_EINVAL;
  qs_ctx *ctx = B->ctx;
  QS_HIP(ctx, hipSetDevice(ctx->device));
  SlotDev S;
  QS_TRY(qs_d2h(ctx, &S, B->d_slots + slot, sizeof(SlotDev)));
  QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (step) *step = S.step;
  const SlotDev &H = B->h_slots[slot];
  if (a) QS_TRY(qs_d2h(ctx, a, H.a, (size_t)(H.depth + 1) * 8));
  if (b) QS_TRY(qs_d2h(ctx, b, H.b, (size_t)(H.depth + 1) * 8));
  if (y) {
    if (!H.y) return qs_fail(ctx, QS_EINVAL, "slot %d accumulates no vector", slot);
    if (H.A.kind == 1) QS_TRY(qs_prod_from_rows(ctx, B->hops[slot], 1, H.y, H.n, (double2 *)y, H.n));
    else QS_HIP(ctx, hipMemcpyAsync(y, H.y, (size_t)H.n * 16, hipMemcpyDeviceToDevice, ctx->stream));
  }
  QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QS_OK;
}

extern "C" qs_status qs_batch_ritz(qs_batch *B, int slot, int m, const double *z, double *y) {
  if (!B || slot < 0 || slot >= B->cap || !B->used[slot] || m < 0 || (m > 0 && (!z || !y))) return QS_EINVAL;
  qs_ctx *ctx = B->ctx;
  const SlotDev &H = B->h_slots[slot];
  if (!H.V) return qs_fail(ctx, QS_EINVAL, "slot %d keeps no Lanczos vectors", slot);
  QS_HIP(ctx, hipSetDevice(ctx->device));
  SlotDev S;
  QS_TRY(qs_d2h(ctx, &S, B->d_slots + slot, sizeof(SlotDev)));
  QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (m > S.step) return qs_fail(ctx, QS_EINVAL, "Ritz vector over %d vectors, slot %d has %d", m, slot, S.step);
  void *scr;
  QS_TRY(qs_scratch(ctx, (size_t)std::max(m, 1) * 8, &scr));
  if (m > 0) QS_TRY(qs_h2d(ctx, scr, z, (size_t)m * 8));
  const int nb = (int)std::max<int64_t>(1, std::min<int64_t>((H.n + QS_BLOCK - 1) / QS_BLOCK, (int64_t)ctx->num_cu * 8));
  const int32_t *inv = H.A.kind == 1 ? B->hops[slot]->d_inv : nullptr;
  hipLaunchKernelGGL(k_b_ritz, dim3(nb), dim3(QS_BLOCK), 0, ctx->stream, B->d_slots, slot, m, (const double *)scr,
                     inv, (double2 *)y);
  QS_HIP(ctx, hipGetLastError());
  QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QS_OK;
}

extern "C" qs_status qs_batch_remove(qs_batch *B, int slot) {
  if (!B || slot < 0 || slot >= B->cap || !B->used[slot]) return QS_EINVAL;
  qs_ctx *ctx = B->ctx;
  // an empty descriptor (active = 0) after the steps in flight; the slot's
  // memory goes back once the stream has passed it
  hipLaunchKernelGGL(k_b_slot_set, dim3(1), dim3(1), 0, ctx->stream, B->d_slots, slot, SlotDev{}, (double *)nullptr, 0);
  QS_HIP(ctx, hipGetLastError());
  QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  pool_put(B, B->mem[slot], B->mem_bytes[slot]);
  B->mem[slot] = nullptr;
  B->mem_bytes[slot] = 0;
  pool_put(B, B->kept[slot], B->kept_bytes[slot]);
  B->kept[slot] = nullptr;
  B->kept_bytes[slot] = 0;
  B->used[slot] = 0;
  B->hops[slot] = nullptr;
  B->h_slots[slot] = SlotDev{};
  B->dirty = true;
  return QS_OK;
}

// one workgroup per listed slot: its step count and a[0..ld), b[0..ld) (zero
// past the slot's capacity) packed into out = [steps (n ints, padded to 8 B) |
// a (n x ld) | b (n x ld)]
__global__ void __launch_bounds__(QS_BLOCK)
k_b_gather(const SlotDev *__restrict__ slots, int n, const int *__restrict__ list, int ld, char *__restrict__ out) {
  const int i = blockIdx.x;
  if (i >= n) return;
  const SlotDev &S = slots[list[i]];
  int *steps = reinterpret_cast<int *>(out);
  double *a = reinterpret_cast<double *>(out + (size_t)((n + 1) / 2) * 8);
  double *b = a + (size_t)n * ld;
  if (threadIdx.x == 0) steps[i] = S.active ? S.step : -1;
  const int cap = S.pad + 1;
  for (int k = threadIdx.x; k < ld; k += blockDim.x) {
    a[(size_t)i * ld + k] = k < cap ? S.a[k] : 0.0;
    b[(size_t)i * ld + k] = k < cap ? S.b[k] : 0.0;
  }
}

extern "C" qs_status qs_batch_coeffs(qs_batch *B, int n, const int *slots, int ld, int *steps, double *a,
                                     double *b) {
  if (!B || n < 0 || ld < 1 || (n > 0 && (!slots || !steps || !a || !b))) return QS_EINVAL;
  qs_ctx *ctx = B->ctx;
  for (int i = 0; i < n; ++i)
    if (slots[i] < 0 || slots[i] >= B->cap || !B->used[slots[i]])
      return qs_fail(ctx, QS_EINVAL, "slot %d is not in use", slots[i]);
  if (n == 0) return QS_OK;
  QS_HIP(ctx, hipSetDevice(ctx->device));
  const size_t head = (size_t)((n + 1) / 2) * 8;
  const size_t body = head + (size_t)2 * n * ld * 8;
  void *scr;
  QS_TRY(qs_scratch(ctx, (size_t)n * 4 + 64 + body, &scr));
  int *d_list = (int *)scr;
  char *d_out = (char *)scr + (((size_t)n * 4 + 63) / 64) * 64;
  QS_TRY(qs_h2d(ctx, d_list, slots, (size_t)n * 4));
  hipLaunchKernelGGL(k_b_gather, dim3(n), dim3(QS_BLOCK), 0, ctx->stream, B->d_slots, n, d_list, ld, d_out);
  QS_HIP(ctx, hipGetLastError());
  std::vector<char> h(body);
  QS_TRY(qs_d2h(ctx, h.data(), d_out, body));
  QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  memcpy(steps, h.data(), (size_t)n * 4);
  memcpy(a, h.data() + head, (size_t)n * ld * 8);
  memcpy(b, h.data() + head + (size_t)n * ld * 8, (size_t)n * ld * 8);
  return QS_OK;
}

__global__ void k_b_set_depths(SlotDev *slots, int n, const int *__restrict__ list) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) slots[list[2 * i]].depth = list[2 * i + 1];
}

extern "C" qs_status qs_batch_set_depths(qs_batch *B, int n, const int *slots, const int *depths) {
  if (!B || n < 0 || (n > 0 && (!slots || !depths))) return QS_EINVAL;
  qs_ctx *ctx = B->ctx;
  for (int i = 0; i < n; ++i) {
    if (slots[i] < 0 || slots[i] >= B->cap || !B->used[slots[i]])
      return qs_fail(ctx, QS_EINVAL, "slot %d is not in use", slots[i]);
    const SlotDev &H = B->h_slots[slots[i]];
    if (depths[i] < 1 || depths[i] > H.pad)
      return qs_fail(ctx, QS_EINVAL, "depth %d outside 1..%d", depths[i], H.pad);
    if (H.y) return qs_fail(ctx, QS_EINVAL, "the depth of a vector-building slot is fixed");
  }
  if (n == 0) return QS_OK;
  QS_HIP(ctx, hipSetDevice(ctx->device));
  std::vector<int> pairs((size_t)2 * n);
  for (int i = 0; i < n; ++i) {
    pairs[2 * i] = slots[i];
    pairs[2 * i + 1] = depths[i];
  }
  void *scr;
  QS_TRY(qs_scratch(ctx, pairs.size() * 4, &scr));
  QS_TRY(qs_h2d(ctx, scr, pairs.data(), pairs.size() * 4));
  hipLaunchKernelGGL(k_b_set_depths, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, B->d_slots, n,
                     (const int *)scr);
  QS_HIP(ctx, hipGetLastError());
  // the host copy of `pairs` must outlive the transfer
  QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < n; ++i) B->h_slots[slots[i]].depth = depths[i];
  return QS_OK;
}

__global__ void k_b_steps(const SlotDev *__restrict__ slots, int n, int *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = slots[i].active ? slots[i].step : -1;
}

extern "C" qs_status qs_batch_steps(qs_batch *B, int *steps) {
  if (!B || !steps) return QS_EINVAL;
  qs_ctx *ctx = B->ctx;
  QS_HIP(ctx, hipSetDevice(ctx->device));
  void *scr;
  QS_TRY(qs_scratch(ctx, (size_t)B->cap * 4, &scr));
  hipLaunchKernelGGL(k_b_steps, dim3((B->cap + 255) / 256), dim3(256), 0, ctx->stream, B->d_slots, B->cap, (int *)scr);
  QS_HIP(ctx, hipGetLastError());
  QS_TRY(qs_d2h(ctx, steps, scr, (size_t)B->cap * 4));
  QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QS_OK;
}

extern "C" int qs_batch_active(const qs_batch *B) {
  if (!B) return -1;
  int k = 0;
  for (int u : B->used) k += u;
  return k;
}

extern "C" qs_status qs_batch_destroy(qs_batch *B) {
  if (!B) return QS_OK;
  if (B->ctx) (void)hipStreamSynchronize(B->ctx->stream);
  for (char *m : B->mem)
    qs_free(m);
  for (double2 *v : B->kept)
    qs_free(v);
  qs_free(B->d_slots);
  qs_free(B->d_w1);
  qs_free(B->d_w2);
  drop_graphs(B);
  qs_free(B->d_n);
  qs_free(B->d_r1);
  qs_free(B->d_r2);
  qs_free(B->d_off);
  qs_free(B->d_p1);
  qs_free(B->d_p2);
  delete B;
  return QS_OK;
}
