// Device building blocks shared by the apply, Lanczos and batch translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "qs_internal.h"

#ifndef QS_BLOCK
#define QS_BLOCK 256
#endif


// Minimum workgroups per CU asked of the register allocator for the fused
// Lanczos kernel's wide variants (two-word determinants or unpacked entries).
#ifndef QS_MINB_WIDE
#define QS_MINB_WIDE 1
#endif

// Wave-uniform copy of a value the compiler cannot prove uniform.
__device__ __forceinline__ int64_t wave_uniform(int64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Element i of read-only data that every lane of a wave reads at the same
// address: loads through the constant address space go to the scalar cache
// (s_load, adjacent words merged) instead
// of returning one copy per lane through the vector memory path (texture
// addresser and data units). Only for buffers no kernel writes while the
// reading kernel runs.
template <typename T>
__device__ __forceinline__ T ld_cst(const T *p, int64_t i) {
  static_assert(sizeof(T) % 4 == 0, "whole 32-bit words");
  typedef const __attribute__((address_space(4))) uint32_t *cword;
  const cword q = (cword)(p + i);
  T out;
  uint32_t *o = reinterpret_cast<uint32_t *>(&out);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / 4); ++k) o[k] = q[k];
  return out;
}

// Position j of lane `lane` in the ELL part of a slice starting at `base`:
// a lane's positions are stored four at a time (quad j/4 of the slice holds
// 64 lanes x 4 consecutive positions), so one 16-byte load per lane fetches
// four positions.
__host__ __device__ __forceinline__ uint64_t ell_slot(uint64_t base, uint32_t j, int lane) {
  return base + ((uint64_t)(j >> 2) << 8) + (uint64_t)lane * 4 + (j & 3u);
}

// Term t applied to the row's determinant d: accumulate
// (-1)^popcount(d & P_t) V_t x[c] for KC interleaved vectors when `hit`.
template <int KC, int NW>
__device__ __forceinline__ void sell_acc(bool hit, int64_t c, double2 v, const uint64_t *P, const uint64_t *d,
                                         const double2 *__restrict__ x, int64_t ldx, int64_t ldv, double *ar,
                                         double *ai) {
  uint32_t p = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) p += __popcll(d[w] & P[w]);
  const uint64_t sb = (uint64_t)(p & 1u) << 63;
  v.x = __longlong_as_double(__double_as_longlong(v.x) ^ sb);
  v.y = __longlong_as_double(__double_as_longlong(v.y) ^ sb);
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    const double2 xv = x[c * ldx + k * ldv];
    const double tr = fma(v.x, xv.x, fma(-v.y, xv.y, ar[k]));
    const double ti = fma(v.x, xv.y, fma(v.y, xv.x, ai[k]));
    ar[k] = hit ? tr : ar[k];
    ai[k] = hit ? ti : ai[k];
  }
}

// Row s*64 + lane of the assembled operator A applied to KC vectors, entry j
// of vector k at x[j*ldx + k*ldv] (interleaved: ldx = KC, ldv = 1; one vector
// after another: ldx = 1, ldv = n). `s` must be wave-uniform.
// tv / tp: the term values and parity masks (LDS copies or A.val / A.par).
// Both loops are branch-free so that many gathers are in flight per wave:
// lanes without an entry read x[0] and keep their sums.
template <int KC, int NW, bool PACKED>
__device__ __forceinline__ void sell_row(const SellD &A, int64_t s, int lane, const double2 *tv,
                                         const uint64_t *tp, const double2 *__restrict__ x, int64_t ldx,
                                         int64_t ldv, double *ar, double *ai) {
  constexpr int U = KC == 1 ? 8 : 4;  // ELL positions per batch
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    ar[k] = 0.0;
    ai[k] = 0.0;
  }
  const int64_t r = s * 64 + lane;
  const bool active = r < A.nrows;
  uint64_t d[NW];
  uint32_t cnt = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) d[w] = 0;
  if (active) {
#pragma unroll
    for (int w = 0; w < NW; ++w) d[w] = A.rows[r * NW + w];
    cnt = A.rowcnt[r];
    if (A.diag) {
      if (A.real_diag) {
        const double dg = static_cast<const double *>(A.diag)[r];
#pragma unroll
        for (int k = 0; k < KC; ++k) {
          const double2 xv = x[r * ldx + k * ldv];
          ar[k] = dg * xv.x;
          ai[k] = dg * xv.y;
        }
      } else {
        const double2 dg = static_cast<const double2 *>(A.diag)[r];
#pragma unroll
        for (int k = 0; k < KC; ++k) {
          const double2 xv = x[r * ldx + k * ldv];
          ar[k] = dg.x * xv.x - dg.y * xv.y;
          ai[k] = dg.x * xv.y + dg.y * xv.x;
        }
      }
    }
  }
  // shift blocks: up to 64 headers are loaded at once (lane i holds header i)
  // and broadcast one by one; the hit lanes of a block read x[base + lane].
  // Groups of 4 headers with a fixed trip count let the compiler overlap their
  // gathers (a loop of readlanes is not unrolled at a runtime trip count);
  // lanes past the last header hold an empty mask.
  const uint64_t b0 = A.blk_ptr[s], b1 = A.blk_ptr[s + 1];
  for (uint64_t c0 = b0; c0 < b1; c0 += 64) {
    const int nb = (int)((b1 - c0) < 64 ? (b1 - c0) : 64);
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    u32x4 hv = {0u, 0u, 0u, 0u};
    if (lane < nb) hv = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(A.blk + c0 + lane));
    for (int k0 = 0; k0 < nb; k0 += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = k0 + u;
        const int32_t base = (int32_t)__builtin_amdgcn_readlane(hv.x, k);
        const uint32_t t = __builtin_amdgcn_readlane(hv.y, k);
        const uint64_t mask = ((uint64_t)__builtin_amdgcn_readlane(hv.w, k) << 32) |
                              (uint32_t)__builtin_amdgcn_readlane(hv.z, k);
        const bool hit = (mask >> lane) & 1u;
        const int64_t c = hit ? (int64_t)base + lane : 0;
        sell_acc<KC, NW>(hit, c, tv[t], tp + (size_t)t * NW, d, x, ldx, ldv, ar, ai);
      }
    }
  }
  // ELL remainder, U positions per batch, software-pipelined: the stream
  // loads of batch b+1 are issued before the gathers of batch b, so the two
  // latencies overlap. A lane's positions come four per 16-byte load (slices
  // hold whole quads). Padding slots hold column 0 / term 0 and are masked.
  // The stream is read once per apply, so non-temporal loads keep it from
  // evicting the gathered vector from L2 / the Infinity Cache.
  typedef uint32_t u32x4q __attribute__((ext_vector_type(4)));
  typedef uint32_t u32x2q __attribute__((ext_vector_type(2)));
  constexpr int NQ = U / 4;
  const uint64_t sb = A.slice_ptr[s];
  const uint32_t slen = A.slice_len[s];
  const uint32_t nq = slen >> 2;
  const u32x4q *cp = reinterpret_cast<const u32x4q *>(A.col + sb) + lane;
  const u32x2q *tq = PACKED ? nullptr : reinterpret_cast<const u32x2q *>(A.tid + sb) + lane;
  const uint32_t cbits = A.colbits;
  const uint32_t cmask = PACKED ? ((1u << cbits) - 1u) : 0xFFFFFFFFu;
  uint32_t e[U], tt[U];
  auto load_batch = [&](uint32_t j0, uint32_t *eo, uint32_t *to) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const uint32_t qi = (j0 >> 2) + q < nq ? (j0 >> 2) + q : nq - 1;
      const u32x4q v = __builtin_nontemporal_load(cp + (size_t)qi * 64);
      eo[4 * q + 0] = v.x;
      eo[4 * q + 1] = v.y;
      eo[4 * q + 2] = v.z;
      eo[4 * q + 3] = v.w;
      if (PACKED) {
        to[4 * q + 0] = to[4 * q + 1] = to[4 * q + 2] = to[4 * q + 3] = 0u;
      } else {
        const u32x2q t = __builtin_nontemporal_load(tq + (size_t)qi * 64);
        to[4 * q + 0] = t.x & 0xFFFFu;
        to[4 * q + 1] = t.x >> 16;
        to[4 * q + 2] = t.y & 0xFFFFu;
        to[4 * q + 3] = t.y >> 16;
      }
    }
  };
  if (slen > 0) load_batch(0, e, tt);
  for (uint32_t j0 = 0; j0 < slen; j0 += U) {
    uint32_t en[U], tn[U];
    const bool more = j0 + U < slen;
    if (more) load_batch(j0 + U, en, tn);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool ok = j0 + u < cnt;
      const uint32_t c = ok ? (e[u] & cmask) : 0u;
      const uint32_t t = ok ? (PACKED ? e[u] >> cbits : tt[u]) : 0u;
      sell_acc<KC, NW>(ok, (int64_t)c, tv[t], tp + (size_t)t * NW, d, x, ldx, ldv, ar, ai);
    }
    if (more) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        e[u] = en[u];
        tt[u] = tn[u];
      }
    }
  }
}

// Rows of one group that a wavefront stages in LDS for the A-matrix (product
// layout, one vector): groups with more rows read their own rows from memory.
#define QS_PROD_STAGE 256
// zero entries stored past the end of a product binding's entry list, so a
// batch of entries is loaded without bounds tests
#define QS_PROD_ENT_PAD 8

// Row of lane `lane` in product slice s (see SellD, kind 1) applied to KC
// vectors (entry j of vector k at x[j*ldx + k*ldv]); returns the row, or -1
// for a lane past the slice's rows (its sums stay 0). `s` must be
// wave-uniform, so the slice header and the group's entries are scalar
// loads. tv: the term values (an LDS copy or A.val). xs: this wavefront's LDS
// area of QS_PROD_STAGE vector entries, or null; with KC == 1 and a group of
// at most QS_PROD_STAGE rows, the wavefront copies the group's rows there and
// the A-matrix reads them from LDS.
//   diagonal: the row's pre-summed diagonal;
//   A-matrix: per row, 4-byte slots (source index in the group | term | sign)
//     over the row's own group;
//   entries: per B-moving term, its source group; lanes of an identity entry
//     read consecutive rows (one contiguous load per entry), the others go
//     through the entry's map (source index, sign in bit 30).
// Lanes past the slice's rows load nothing.
template <int KC>
__device__ __forceinline__ int64_t prod_row(const SellD &A, int64_t s, int lane, const double2 *tv,
                                            const double2 *__restrict__ x, int64_t ldx, int64_t ldv, double *ar,
                                            double *ai, double2 *xs) {
  constexpr int U = KC == 1 ? 8 : 4;   // entries per batch
  constexpr int UA = 8;                // A-matrix slots per batch
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    ar[k] = 0.0;
    ai[k] = 0.0;
  }
  s = wave_uniform(s);
  // the slice, its A-set and its group's entries are the same for every lane:
  // scalar loads
  const int4 d = ld_cst(A.pslice, s);
  const bool active = lane < d.y;
  const int chunk = d.w >> 16;
  const int ap = chunk * 64 + lane;                 // index of the row's A configuration in its A-set
  const int64_t g0 = (int64_t)d.x - chunk * 64;     // first row of the group
  const int64_t r = (int64_t)d.x + lane;
  const int4 as = ld_cst(A.aset, d.w & 0xFFFF);    // {m, k, offset, 0}
  const uint2 pp = ld_cst(A.pent_ptr, d.z);
  const uint32_t e0 = pp.x, em = pp.y, e1 = ld_cst(A.pent_ptr, (int64_t)d.z + 1).x;
  const bool staged = KC == 1 && xs != nullptr && as.x <= QS_PROD_STAGE;
  if (staged) {
    // the previous slice's reads of xs are done before it is overwritten
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int i = lane; i < as.x; i += 64) xs[i] = x[(g0 + i) * ldx];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if (!active) return -1;
  if (A.diag) {
    double2 dg;
    if (A.real_diag) dg = make_double2(static_cast<const double *>(A.diag)[r], 0.0);
    else dg = static_cast<const double2 *>(A.diag)[r];
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      const double2 xv = staged ? xs[ap] : x[r * ldx + k * ldv];
      ar[k] = dg.x * xv.x - dg.y * xv.y;
      ai[k] = dg.x * xv.y + dg.y * xv.x;
    }
  }
  // A-matrix, UA slots at a time (one 4-byte load per slot and lane)
  for (int eb = 0; eb < as.y; eb += UA) {
    uint32_t sl[UA];
#pragma unroll
    for (int u = 0; u < UA; ++u) {
      const int e = eb + u < as.y ? eb + u : as.y - 1;
      sl[u] = A.hA[(int64_t)as.z + (int64_t)e * as.x + ap];
    }
#pragma unroll
    for (int u = 0; u < UA; ++u) {
      const uint32_t src = sl[u] & 0xFFFFu;
      const bool hit = eb + u < as.y && src != 0xFFFFu;
      double2 v = tv[(sl[u] >> 16) & 0x7FFFu];
      const uint64_t sb = (uint64_t)(sl[u] >> 31) << 63;
      v.x = __longlong_as_double(__double_as_longlong(v.x) ^ sb);
      v.y = __longlong_as_double(__double_as_longlong(v.y) ^ sb);
      if (!hit) v = make_double2(0.0, 0.0);
      const int64_t c = hit ? g0 + src : r;
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        const double2 xv = staged ? xs[hit ? src : ap] : x[c * ldx + k * ldv];
        ar[k] = fma(v.x, xv.x, fma(-v.y, xv.y, ar[k]));
        ai[k] = fma(v.x, xv.y, fma(v.y, xv.x, ai[k]));
      }
    }
  }
  // identity entries: one coefficient per entry (its terms summed), one
  // contiguous load of the source group's rows
  // (the U entries of a batch are loaded unconditionally, as scalar loads
  // issued together: the list has QS_PROD_ENT_PAD >= U entries of padding)
  static_assert(U <= QS_PROD_ENT_PAD, "entry batch beyond the padding");
  for (uint32_t eb = e0; eb < em; eb += U) {
    int64_t col[U];
    double2 v[U];
    uint2 E[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint4 e4 = ld_cst(A.pent, (int64_t)eb + u);
      E[u] = make_uint2(e4.x, e4.y);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool ok = eb + u < em;
      double2 vv = tv[E[u].y & 0xFFFFu];
      const uint64_t sb = (uint64_t)(E[u].y >> 31) << 63;
      vv.x = __longlong_as_double(__double_as_longlong(vv.x) ^ sb);
      vv.y = __longlong_as_double(__double_as_longlong(vv.y) ^ sb);
      v[u] = ok ? vv : make_double2(0.0, 0.0);
      col[u] = ok ? (int64_t)E[u].x + ap : r;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        const double2 xv = x[col[u] * ldx + k * ldv];
        ar[k] = fma(v[u].x, xv.x, fma(-v[u].y, xv.y, ar[k]));
        ai[k] = fma(v[u].x, xv.y, fma(v[u].y, xv.x, ai[k]));
      }
    }
  }
  // mapped entries: a source index and sign per lane from the map
  constexpr int UM = 4;
  for (uint32_t eb = em; eb < e1; eb += UM) {
    uint4 E[UM];
    int32_t q[UM];
#pragma unroll
    for (int u = 0; u < UM; ++u) E[u] = ld_cst(A.pent, (int64_t)eb + u);  // padded list, as above
#pragma unroll
    for (int u = 0; u < UM; ++u) q[u] = A.amap[eb + u < e1 ? E[u].z + ap : 0];
    int64_t col[UM];
    double2 v[UM];
#pragma unroll
    for (int u = 0; u < UM; ++u) {
      const bool h = eb + u < e1 && q[u] >= 0;
      double2 vv = tv[E[u].y & 0xFFFFu];
      const uint64_t sb = (uint64_t)((E[u].y >> 31) ^ (((uint32_t)q[u] >> 30) & 1u)) << 63;
      vv.x = __longlong_as_double(__double_as_longlong(vv.x) ^ sb);
      vv.y = __longlong_as_double(__double_as_longlong(vv.y) ^ sb);
      v[u] = h ? vv : make_double2(0.0, 0.0);
      col[u] = h ? (int64_t)E[u].x + (q[u] & 0x3FFFFFFF) : r;
    }
#pragma unroll
    for (int u = 0; u < UM; ++u) {
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        const double2 xv = x[col[u] * ldx + k * ldv];
        ar[k] = fma(v[u].x, xv.x, fma(-v[u].y, xv.y, ar[k]));
        ai[k] = fma(v[u].x, xv.y, fma(v[u].y, xv.x, ai[k]));
      }
    }
  }
  return r;
}

// Dynamic LDS of a product-layout kernel: the term values when `values`,
// then one staging area per wavefront (KC == 1 only).
__host__ __device__ inline size_t prod_lds_bytes(int nterms, bool values, int kc) {
  const size_t v = values ? ((size_t)nterms * 16 + 15) / 16 * 16 : 0;
  return v + (kc == 1 ? (size_t)(QS_BLOCK / 64) * QS_PROD_STAGE * 16 : 0);
}

// This wavefront's staging area in a product-layout kernel's dynamic LDS.
__device__ __forceinline__ double2 *prod_stage_area(char *smem, int nterms, bool values, int kc) {
  if (kc != 1) return nullptr;
  const size_t v = values ? ((size_t)nterms * 16 + 15) / 16 * 16 : 0;
  return reinterpret_cast<double2 *>(smem + v) + (size_t)(threadIdx.x >> 6) * QS_PROD_STAGE;
}

// Bytes of LDS that stage_term_values uses.
__host__ __device__ inline size_t term_value_bytes(int nterms) { return (size_t)nterms * 16; }

// Stage the term values of A into LDS at smem (no barrier).
__device__ __forceinline__ void stage_term_values(const SellD &A, char *smem) {
  double2 *sv = reinterpret_cast<double2 *>(smem);
  for (int i = threadIdx.x; i < A.nterms; i += blockDim.x) sv[i] = A.val[i];
}

// Bytes of LDS that stage_term_tables uses.
__host__ __device__ inline size_t term_table_bytes(int nterms, int nw) { return (size_t)nterms * (16 + 8 * nw); }

// Stage the term tables of A into LDS at smem (no barrier).
template <int NW>
__device__ __forceinline__ void stage_term_tables(const SellD &A, char *smem) {
  double2 *sv = reinterpret_cast<double2 *>(smem);
  uint64_t *sp = reinterpret_cast<uint64_t *>(smem + (size_t)A.nterms * 16);
  for (int i = threadIdx.x; i < A.nterms; i += blockDim.x) sv[i] = A.val[i];
  for (int i = threadIdx.x; i < A.nterms * NW; i += blockDim.x) sp[i] = A.par[i];
}

// XCD-aware split of n wave-sized work items. Workgroups are dispatched
// round-robin over the 8 XCDs (workgroup b runs on XCD b % 8), so the items are
// cut into 8 contiguous ranges and the waves of XCD x stride over range x:
// slices that are neighbours in the basis then share one XCD's L2. Grids of
// fewer than 8 workgroups fall back to a plain grid stride.
__device__ __forceinline__ void xcd_items(int64_t n, int64_t &first, int64_t &end, int64_t &stride) {
  constexpr int NX = 8;
  const int wpb = blockDim.x >> 6;
  const int64_t blocks = gridDim.x;
  if (blocks < NX) {
    first = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    end = n;
    stride = (blocks * blockDim.x) >> 6;
    return;
  }
  const int xcd = blockIdx.x % NX;
  const int64_t local = blockIdx.x / NX;
  const int64_t in_xcd = (blocks - xcd + NX - 1) / NX;
  const int64_t b = n * xcd / NX;
  end = n * (xcd + 1) / NX;
  first = b + local * wpb + (threadIdx.x >> 6);
  stride = in_xcd * wpb;
}

// Sum over a wavefront by an xor butterfly: every lane ends with the same
// value, in the same order on every call.
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Lane exchanges without LDS on the two halves of a double: DPP row
// operations and scalar reads of one lane. join_f64: a double from its halves.
__device__ __forceinline__ double join_f64(unsigned lo, unsigned hi) {
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// value of the lane that DPP control CTRL names (row_mirror 0x140: 15 - l in
// a row of 16; row_half_mirror 0x141: 7 - l in 8; quad_perm 0xB1: l xor 1;
// 0x4E: l xor 2)
template <int CTRL> __device__ __forceinline__ double dpp_f64(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
  return join_f64((unsigned)lo, (unsigned)hi);
}
// the value in lane l (the same in every lane)
__device__ __forceinline__ double readlane_f64(double v, int l) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  return join_f64((unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l),
                  (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l));
}
// sum over the 8 lanes of a half-row, the same in each of them
__device__ __forceinline__ double sum8_lanes(double x) {
  x += dpp_f64<0xB1>(x);
  x += dpp_f64<0x4E>(x);
  x += dpp_f64<0x141>(x);
  return x;
}

// Sum over a wavefront in a fixed order, the same in every lane: DPP within
// rows of 16 lanes, then the four row sums in row order; whole wave active.
__device__ __forceinline__ double wave_total(double x) {
  x = sum8_lanes(x);
  x += dpp_f64<0x140>(x);
  return (readlane_f64(x, 0) + readlane_f64(x, 16)) + (readlane_f64(x, 32) + readlane_f64(x, 48));
}

// Block-wide sum of one value per thread in a fixed order; the result is valid
// in every thread. red: at least blockDim.x / 64 doubles of LDS.
__device__ __forceinline__ double block_sum_values(double v, double *red) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  const int nwv = blockDim.x >> 6;
  for (int i = 0; i < nwv; ++i) r += red[i];
  __syncthreads();
  return r;
}

// Sum of `np` partials in a fixed order, identical in every block.
__device__ __forceinline__ double block_sum_fixed(const double *__restrict__ part, int np, double *red) {
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) s += part[i];
  return block_sum_values(s, red);
}
