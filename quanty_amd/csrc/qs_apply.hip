// Operator compilation and H*psi on a fixed basis, in pull form:
//   y_i = sum_t V_t <d_i| t |d_j> x_j,   d_j = t^+ d_i
// computed row by row without atomics. Two execution paths share one hit
// enumerator: the matrix-free kernel recomputes every (row, term) hit on each
// call; the assembled path stores each row's off-diagonal hits once as a
// sliced-ELL stream (64 rows per slice, column-major inside the slice) of
// column index + signed term id, pre-sums the diagonal terms per row, and then
// applies the operator as a streaming gather.
#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>

#include <mutex>
#include "qs_internal.h"
#include "qs_kernels.h"

// ---------------------------------------------------------------------------
// term compilation (host)
// ---------------------------------------------------------------------------
namespace {

struct ModeBits {
  int nw;
  void set(uint64_t *m, int i) const { m[i >> 6] |= 1ULL << (63 - (i & 63)); }
  // modes with index < i
  void low(uint64_t *m, int i) const {
    for (int w = 0; w < QS_MAX_WORDS; ++w) m[w] = 0;
    int wi = i >> 6, b = i & 63;
    for (int w = 0; w < wi; ++w) m[w] = ~0ULL;
    if (b > 0) m[wi] = ~0ULL << (64 - b);
  }
};

int parity_and(const uint64_t *a, const uint64_t *b, int nw) {
  int p = 0;
  for (int w = 0; w < nw; ++w) p ^= __builtin_popcountll(a[w] & b[w]) & 1;
  return p;
}

// Follows the action of OperatorLader (Core/OperatorWaveFunction.cpp:98-116):
// o[0] is applied first; each c+/c of mode q contributes (-1)^(occupied modes
// below q) in the intermediate determinant (OperatorCr/OperatorAn, :7-96).
// Returns false when the string vanishes on every determinant.
bool compile_term(int nf, int len, const uint16_t *o, double re, double im, TermHost &t,
                  std::string &err) {
  const int nw = nf <= 64 ? 1 : 2;
  ModeBits mb{nw};
  std::map<int, std::pair<int, int>> st;  // mode -> (initial occupancy, current occupancy)
  uint64_t G[QS_MAX_WORDS] = {0, 0};
  t = TermHost();
  int c = 0;
  for (int k = 0; k < len; ++k) {
    const bool cr = (o[k] & 0x8000u) != 0;
    const int q = o[k] & 0x7FFF;
    if (q >= nf) {
      err = "ladder operator index beyond the number of fermion modes (bosons unsupported)";
      return false;
    }
    auto it = st.find(q);
    if (it == st.end()) {
      st[q] = {cr ? 0 : 1, cr ? 0 : 1};
      it = st.find(q);
    }
    if (it->second.second != (cr ? 0 : 1)) return false;  // Pauli: always zero
    uint64_t L[QS_MAX_WORDS];
    mb.low(L, q);
    for (int w = 0; w < nw; ++w) t.P[w] ^= L[w];
    c ^= parity_and(G, L, nw);
    it->second.second ^= 1;
    uint64_t bit[QS_MAX_WORDS] = {0, 0};
    mb.set(bit, q);
    for (int w = 0; w < nw; ++w) G[w] ^= bit[w];
    t.nchange += cr ? 1 : -1;
  }
  for (auto &kv : st) {
    mb.set(t.C, kv.first);
    if (kv.second.second) mb.set(t.V, kv.first);
  }
  for (int w = 0; w < nw; ++w) t.F[w] = G[w];
  c ^= parity_and(t.F, t.P, nw);
  t.sgn = (uint32_t)c;
  t.re = re;
  t.im = im;
  return true;
}

}  // namespace

extern "C" qs_status qs_op_create(qs_ctx *ctx, int nf, int nterms, const int *len,
                                  const uint16_t *codes, const double *values, qs_op **out) {
  if (!ctx || !out || nf <= 0 || nf > 128 || nterms < 0) return QS_EINVAL;
  *out = nullptr;
  const int nw = nf <= 64 ? 1 : 2;
  // compile, then merge terms with identical masks (sign folded into the value)
  std::map<std::tuple<uint64_t, uint64_t, uint64_t, uint64_t, uint64_t, uint64_t, uint64_t, uint64_t>,
           size_t>
      seen;
  std::vector<TermHost> terms;
  size_t off = 0;
  for (int t = 0; t < nterms; ++t) {
    if (len[t] < 0) return qs_fail(ctx, QS_EINVAL, "term %d has negative length", t);
    TermHost th;
    std::string err;
    bool ok = compile_term(nf, len[t], codes + off, values[2 * t], values[2 * t + 1], th, err);
    off += (size_t)len[t];
    if (!err.empty()) return qs_fail(ctx, QS_EINVAL, "term %d: %s", t, err.c_str());
    if (!ok) continue;
    if (th.sgn) {
      th.re = -th.re;
      th.im = -th.im;
      th.sgn = 0;
    }
    auto key = std::make_tuple(th.C[0], th.C[1], th.V[0], th.V[1], th.F[0], th.F[1], th.P[0], th.P[1]);
    auto it = seen.find(key);
    if (it == seen.end()) {
      seen[key] = terms.size();
      terms.push_back(th);
    } else {
      terms[it->second].re += th.re;
      terms[it->second].im += th.im;
    }
  }
  std::vector<TermHost> kept;
  for (auto &t : terms)
    if (t.re != 0.0 || t.im != 0.0) kept.push_back(t);
  QS_HIP(ctx, hipSetDevice(ctx->device));
  qs_op *op = new qs_op();
  op->nf = nf;
  op->nw = nw;
  op->nterms = (int)kept.size();
  op->h_terms = kept;
  const size_t tsz = nw == 1 ? sizeof(TermD<1>) : sizeof(TermD<2>);
  std::vector<char> buf(std::max<size_t>(1, kept.size()) * tsz, 0);
  std::vector<double> val(std::max<size_t>(1, kept.size()) * 2, 0.0);
  std::vector<uint64_t> par(std::max<size_t>(1, kept.size()) * nw, 0);
  for (size_t i = 0; i < kept.size(); ++i) {
    const TermHost &h = kept[i];
    const bool diag = (h.F[0] | h.F[1]) == 0;
    if (nw == 1) {
      TermD<1> d;
      d.C[0] = h.C[0]; d.V[0] = h.V[0]; d.F[0] = h.F[0]; d.P[0] = h.P[0];
      d.re = h.re; d.im = h.im; d.sgn = h.sgn; d.diag = diag;
      memcpy(buf.data() + i * tsz, &d, sizeof d);
    } else {
      TermD<2> d;
      for (int w = 0; w < 2; ++w) { d.C[w] = h.C[w]; d.V[w] = h.V[w]; d.F[w] = h.F[w]; d.P[w] = h.P[w]; }
      d.re = h.re; d.im = h.im; d.sgn = h.sgn; d.diag = diag;
      memcpy(buf.data() + i * tsz, &d, sizeof d);
    }
    val[2 * i + 0] = h.re;  // sign constant already folded in (sgn == 0)
    val[2 * i + 1] = h.im;
    for (int w = 0; w < nw; ++w) par[(size_t)i * nw + w] = h.P[w];
  }
  qs_status st = qs_malloc(ctx, &op->d_terms, buf.size());
  if (st == QS_OK) st = qs_malloc(ctx, (void **)&op->d_val, val.size() * 8);
  if (st == QS_OK) st = qs_malloc(ctx, (void **)&op->d_par, par.size() * 8);
  if (st != QS_OK) {
    qs_op_destroy(op);
    return st;
  }
  st = qs_h2d(ctx, op->d_terms, buf.data(), buf.size());
  if (st == QS_OK) st = qs_h2d(ctx, op->d_val, val.data(), val.size() * 8);
  if (st == QS_OK) st = qs_h2d(ctx, op->d_par, par.data(), par.size() * 8);
  if (st != QS_OK) {
    qs_op_destroy(op);
    return st;
  }
  *out = op;
  return QS_OK;
}

extern "C" int qs_op_nterms(const qs_op *op) { return op ? op->nterms : -1; }

extern "C" qs_status qs_op_destroy(qs_op *op) {
  if (!op) return QS_OK;
  qs_free(op->d_terms);
  qs_free(op->d_val);
  qs_free(op->d_par);
  delete op;
  return QS_OK;
}

// ---------------------------------------------------------------------------
// hit enumeration (device)
// ---------------------------------------------------------------------------
#define QS_TILE 256   // terms staged in LDS per pass
#define QS_BLOCK 256  // threads per block; 4 wavefronts = 4 slices of 64 rows

struct Lookup {
  const uint2 *tbl;      // rank table of the column basis, or null
  uint32_t smax;
  int nch;
  const uint64_t *words; // sorted column determinants (binary search when tbl is null)
  int64_t n;
};

template <int NW>
__device__ __forceinline__ int64_t col_lookup(const uint64_t *s, const Lookup &L) {
  if (L.tbl) {
    uint32_t st = 0, r = 0;
    for (int c = 0; c < L.nch; ++c) {
      uint32_t v = (uint32_t)(s[c >> 3] >> (56 - 8 * (c & 7))) & 0xFFu;
      uint2 e = L.tbl[((size_t)c * L.smax + st) * 256 + v];
      if (e.x == QS_DEAD) return -1;
      r += e.y;
      st = e.x;
    }
    return r;
  }
  int64_t lo = 0, hi = L.n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    const uint64_t *m = L.words + mid * NW;
    bool lt = NW == 1 ? (m[0] < s[0]) : (m[0] < s[0] || (m[0] == s[0] && m[NW - 1] < s[NW - 1]));
    if (lt) lo = mid + 1; else hi = mid;
  }
  if (lo < L.n) {
    bool eq = true;
    for (int w = 0; w < NW; ++w) eq = eq && (L.words[lo * NW + w] == s[w]);
    if (eq) return lo;
  }
  return -1;
}

// Stage terms [base, base+cnt) into LDS.
template <int NW>
__device__ __forceinline__ void stage_terms(TermD<NW> *sh, const TermD<NW> *__restrict__ g, int base,
                                            int cnt) {
  const int nwords = cnt * (int)(sizeof(TermD<NW>) / 8);
  const uint64_t *src = reinterpret_cast<const uint64_t *>(g + base);
  uint64_t *dst = reinterpret_cast<uint64_t *>(sh);
  for (int i = threadIdx.x; i < nwords; i += blockDim.x) dst[i] = src[i];
}

// Visit every hit of row `row` (determinant d) for the staged terms.
// f(term_index, column, parity, is_diagonal)
template <int NW, class F>
__device__ __forceinline__ void visit_tile(const uint64_t *d, int64_t row, const TermD<NW> *sh, int cnt,
                                           int tbase, bool same, const Lookup &L, F &&f) {
  for (int j = 0; j < cnt; ++j) {
    const TermD<NW> &t = sh[j];
    bool hit = true;
#pragma unroll
    for (int w = 0; w < NW; ++w) hit = hit && ((d[w] & t.C[w]) == t.V[w]);
    if (!hit) continue;
    const bool dg = t.diag && same;
    int64_t col;
    if (dg) {
      col = row;
    } else {
      uint64_t s[NW];
#pragma unroll
      for (int w = 0; w < NW; ++w) s[w] = d[w] ^ t.F[w];
      col = col_lookup<NW>(s, L);
    }
    if (col < 0) continue;
    uint32_t par = t.sgn;
#pragma unroll
    for (int w = 0; w < NW; ++w) par ^= (uint32_t)__popcll(d[w] & t.P[w]) & 1u;
    f(tbase + j, col, par, dg);
  }
}

// ---------------------------------------------------------------------------
// matrix-free apply: one lane per row
// ---------------------------------------------------------------------------
template <int NW>
__global__ void __launch_bounds__(QS_BLOCK)
k_apply_mf(const uint64_t *__restrict__ rows, int64_t nrows, const TermD<NW> *__restrict__ terms,
           int nterms, Lookup L, int same, const double2 *__restrict__ x, double2 *__restrict__ y) {
  __shared__ TermD<NW> sh[QS_TILE];
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = row < nrows;
  uint64_t d[NW];
#pragma unroll
  for (int w = 0; w < NW; ++w) d[w] = active ? rows[row * NW + w] : 0;
  double ar = 0.0, ai = 0.0;
  for (int base = 0; base < nterms; base += QS_TILE) {
    const int cnt = min(QS_TILE, nterms - base);
    __syncthreads();
    stage_terms<NW>(sh, terms, base, cnt);
    __syncthreads();
    if (active) {
      visit_tile<NW>(d, row, sh, cnt, base, same != 0, L, [&](int tj, int64_t col, uint32_t par, bool) {
        const TermD<NW> &t = sh[tj - base];
        const double2 xv = x[col];
        double cr = par ? -t.re : t.re, ci = par ? -t.im : t.im;
        ar = fma(cr, xv.x, fma(-ci, xv.y, ar));
        ai = fma(cr, xv.y, fma(ci, xv.x, ai));
      });
    }
  }
  if (active) y[row] = make_double2(ar, ai);
}

// Small bases: g lanes per row (g a power of two up to 64), lane k of a row's
// group taking the terms j = k (mod g) of each tile, then an xor butterfly
// over the group. Keeps enough threads in flight when rows are few and every
// row walks hundreds of terms.
template <int NW>
__global__ void __launch_bounds__(QS_BLOCK)
k_apply_mf_split(const uint64_t *__restrict__ rows, int64_t nrows, const TermD<NW> *__restrict__ terms,
                 int nterms, Lookup L, int same, int g, const double2 *__restrict__ x, double2 *__restrict__ y) {
  __shared__ TermD<NW> sh[QS_TILE];
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / g;
  const int k = threadIdx.x & (g - 1);
  const bool active = row < nrows;
  uint64_t d[NW];
#pragma unroll
  for (int w = 0; w < NW; ++w) d[w] = active ? rows[row * NW + w] : 0;
  double ar = 0.0, ai = 0.0;
  for (int base = 0; base < nterms; base += QS_TILE) {
    const int cnt = min(QS_TILE, nterms - base);
    __syncthreads();
    stage_terms<NW>(sh, terms, base, cnt);
    __syncthreads();
    if (!active) continue;
    for (int j = k; j < cnt; j += g) {
      const TermD<NW> &t = sh[j];
      bool hit = true;
#pragma unroll
      for (int w = 0; w < NW; ++w) hit = hit && ((d[w] & t.C[w]) == t.V[w]);
      if (!hit) continue;
      int64_t col;
      if (t.diag && same) {
        col = row;
      } else {
        uint64_t sv[NW];
#pragma unroll
        for (int w = 0; w < NW; ++w) sv[w] = d[w] ^ t.F[w];
        col = col_lookup<NW>(sv, L);
      }
      if (col < 0) continue;
      uint32_t par = t.sgn;
#pragma unroll
      for (int w = 0; w < NW; ++w) par ^= (uint32_t)__popcll(d[w] & t.P[w]) & 1u;
      const double2 xv = x[col];
      const double cr = par ? -t.re : t.re, ci = par ? -t.im : t.im;
      ar = fma(cr, xv.x, fma(-ci, xv.y, ar));
      ai = fma(cr, xv.y, fma(ci, xv.x, ai));
    }
  }
  for (int o = g >> 1; o > 0; o >>= 1) {
    ar += __shfl_xor(ar, o);
    ai += __shfl_xor(ai, o);
  }
  if (active && k == 0) y[row] = make_double2(ar, ai);
}

// ---------------------------------------------------------------------------
// assembly: one wavefront per 64-row slice, count pass then fill pass
// ---------------------------------------------------------------------------
// For every term the wave finds the hit lanes and their columns, then groups
// the hit lanes by shift (column - lane): a group of at least `hmin` lanes
// becomes one shift block, the lanes of smaller groups get ELL entries. Both
// passes visit terms and groups in the same order, so the fill pass writes
// exactly the layout the count pass sized. Count outputs: ELL entries per row,
// ELL positions per slice, blocks per slice, the pre-summed diagonal and the
// counters [hits, off-diagonal hits, block hits].
template <int NW, bool FILL>
__global__ void __launch_bounds__(QS_BLOCK)
k_sell_build(const uint64_t *__restrict__ rows, int64_t nrows, const TermD<NW> *__restrict__ terms, int nterms,
             Lookup L, int same, int hmin, uint16_t *__restrict__ rowcnt, uint32_t *__restrict__ slice_len,
             uint32_t *__restrict__ blk_cnt, void *__restrict__ diag, int real_diag, unsigned long long *__restrict__ ctr,
             const uint64_t *__restrict__ blk_ptr, uint4 *__restrict__ blk, const uint64_t *__restrict__ slice_ptr,
             uint32_t *__restrict__ colv, uint16_t *__restrict__ tidv, int packed, uint32_t colbits) {
  __shared__ TermD<NW> sh[QS_TILE];
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = row < nrows;
  const int lane = threadIdx.x & 63;
  const int64_t slice = row >> 6;
  const bool live = slice * 64 < nrows;  // wave-uniform
  uint64_t d[NW];
#pragma unroll
  for (int w = 0; w < NW; ++w) d[w] = active ? rows[row * NW + w] : 0;
  uint64_t bbase = 0, ebase = 0;
  if (FILL && live) {
    bbase = blk_ptr[slice];
    ebase = slice_ptr[slice];
  }
  uint32_t nell = 0, nblk = 0;
  unsigned long long eff = 0, bhits = 0;
  double dr = 0.0, di = 0.0;
  for (int base = 0; base < nterms; base += QS_TILE) {
    const int cnt = min(QS_TILE, nterms - base);
    __syncthreads();
    stage_terms<NW>(sh, terms, base, cnt);
    __syncthreads();
    if (!live) continue;
    for (int j = 0; j < cnt; ++j) {
      const TermD<NW> &t = sh[j];
      bool hit = active;
#pragma unroll
      for (int w = 0; w < NW; ++w) hit = hit && ((d[w] & t.C[w]) == t.V[w]);
      if (__ballot(hit) == 0) continue;
      if (t.diag && same) {
        if (hit) {
          uint32_t p = 0;
#pragma unroll
          for (int w = 0; w < NW; ++w) p += __popcll(d[w] & t.P[w]);
          ++eff;
          dr += (p & 1u) ? -t.re : t.re;
          di += (p & 1u) ? -t.im : t.im;
        }
        continue;
      }
      int64_t col = -1;
      if (hit) {
        uint64_t sdet[NW];
#pragma unroll
        for (int w = 0; w < NW; ++w) sdet[w] = d[w] ^ t.F[w];
        col = col_lookup<NW>(sdet, L);
      }
      hit = col >= 0;
      eff += hit ? 1 : 0;
      const int off = hit ? (int)(col - lane) : 0;
      uint64_t un = __ballot(hit);
      const uint32_t tj = (uint32_t)(base + j);
      while (un) {
        const int leader = __ffsll((unsigned long long)un) - 1;
        const int cand = __shfl(off, leader);
        const uint64_t grp = __ballot(hit && off == cand);
        const int ng = __popcll(grp);
        if (ng >= hmin) {
          if (FILL && lane == 0)
            blk[bbase + nblk] = make_uint4((uint32_t)cand, tj, (uint32_t)grp, (uint32_t)(grp >> 32));
          ++nblk;
          bhits += (unsigned long long)ng;
        } else if ((grp >> lane) & 1u) {
          if (FILL) {
            const uint64_t e = ell_slot(ebase, nell, lane);
            if (packed) {
              colv[e] = (uint32_t)col | (tj << colbits);
            } else {
              colv[e] = (uint32_t)col;
              tidv[e] = (uint16_t)tj;
            }
          }
          ++nell;
        }
        un &= ~grp;
      }
    }
  }
  if (!live) return;
  if (FILL) {
    // padding slots of the lane's ELL column: column 0, term 0
    const uint32_t slen = slice_len[slice];
    for (uint32_t j = nell; j < slen; ++j) {
      const uint64_t e = ell_slot(ebase, j, lane);
      colv[e] = 0;
      if (!packed) tidv[e] = 0;
    }
    return;
  }
  if (active) {
    rowcnt[row] = (uint16_t)nell;
    if (diag) {
      if (real_diag) static_cast<double *>(diag)[row] = dr;
      else static_cast<double2 *>(diag)[row] = make_double2(dr, di);
    }
  }
  uint32_t m = nell;
  unsigned long long e = eff;
  for (int o = 32; o > 0; o >>= 1) {
    m = max(m, (uint32_t)__shfl_xor((int)m, o));
    e += (unsigned long long)__shfl_xor((long long)e, o);
  }
  if (lane == 0) {
    slice_len[slice] = (m + 3u) & ~3u;  // whole quads of positions
    blk_cnt[slice] = nblk;
    atomicAdd(ctr, e);
    atomicAdd(ctr + 2, bhits);
  }
  // off-diagonal hits: ELL entries plus block hits
  unsigned long long ne = nell;
  for (int o = 32; o > 0; o >>= 1) ne += (unsigned long long)__shfl_xor((long long)ne, o);
  if (lane == 0) atomicAdd(ctr + 1, ne + bhits);
}

// Reorder every row's ELL entries by column (one thread per row; rows with
// more than QS_SORT_MAX entries keep their order). The padding slots after a
// row's count stay where they are.
#define QS_SORT_MAX 96
template <bool PACKED>
__global__ void __launch_bounds__(QS_BLOCK)
k_sell_sort_rows(int64_t nrows, const uint16_t *__restrict__ rowcnt, const uint64_t *__restrict__ slice_ptr,
                 uint32_t colbits, uint32_t *__restrict__ colv, uint16_t *__restrict__ tidv) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= nrows) return;
  const int cnt = rowcnt[row];
  if (cnt < 2 || cnt > QS_SORT_MAX) return;
  const uint64_t base = slice_ptr[row >> 6];
  const int lane = (int)(row & 63);
  const uint32_t cmask = PACKED ? ((1u << colbits) - 1u) : 0xFFFFFFFFu;
  uint32_t ev[QS_SORT_MAX];
  uint16_t tv[QS_SORT_MAX];
  for (int j = 0; j < cnt; ++j) {
    ev[j] = colv[ell_slot(base, (uint32_t)j, lane)];
    tv[j] = PACKED ? 0 : tidv[ell_slot(base, (uint32_t)j, lane)];
  }
  for (int j = 1; j < cnt; ++j) {  // insertion sort by column
    const uint32_t e = ev[j];
    const uint16_t t = tv[j];
    int k = j - 1;
    while (k >= 0 && (ev[k] & cmask) > (e & cmask)) {
      ev[k + 1] = ev[k];
      tv[k + 1] = tv[k];
      --k;
    }
    ev[k + 1] = e;
    tv[k + 1] = t;
  }
  for (int j = 0; j < cnt; ++j) {
    colv[ell_slot(base, (uint32_t)j, lane)] = ev[j];
    if (!PACKED) tidv[ell_slot(base, (uint32_t)j, lane)] = tv[j];
  }
}

// ---------------------------------------------------------------------------
// assembled apply: one wavefront per slice, lane per row
// ---------------------------------------------------------------------------
template <int KC, int NW, bool PACKED, bool LDS>
__global__ void __launch_bounds__(QS_BLOCK)
k_spmv(SellD A, const double2 *__restrict__ x, int64_t ldx, int64_t ldvx, double2 *__restrict__ y, int64_t ldy,
       int64_t ldvy) {
  extern __shared__ char smem[];
  const double2 *tv = A.val;
  const uint64_t *tp = A.par;
  if (LDS) {
    stage_term_tables<NW>(A, smem);
    __syncthreads();
    tv = reinterpret_cast<const double2 *>(smem);
    tp = reinterpret_cast<const uint64_t *>(smem + (size_t)A.nterms * 16);
  }
  const int lane = threadIdx.x & 63;
  int64_t s0, s1, ds;
  xcd_items(A.nslices, s0, s1, ds);
  s0 = wave_uniform(s0);
  for (int64_t s = s0; s < s1; s += ds) {
    double ar[KC], ai[KC];
    sell_row<KC, NW, PACKED>(A, s, lane, tv, tp, x, ldx, ldvx, ar, ai);
    const int64_t r = s * 64 + lane;
    if (r < A.nrows) {
#pragma unroll
      for (int k = 0; k < KC; ++k) y[r * ldy + k * ldvy] = make_double2(ar[k], ai[k]);
    }
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static Lookup make_lookup(const qs_basis *cols) {
  Lookup L;
  L.tbl = cols->has_rank ? cols->d_tbl : nullptr;
  L.smax = cols->smax;
  L.nch = cols->nchunks;
  L.words = cols->d_words;
  L.n = cols->n;
  return L;
}

static qs_status check_pair(qs_ctx *ctx, const qs_op *op, const qs_basis *rows, const qs_basis *cols) {
  if (!op || !rows || !cols) return qs_fail(ctx, QS_EINVAL, "null operator or basis");
  if (op->nf != rows->nf || op->nf != cols->nf)
    return qs_fail(ctx, QS_EINVAL, "operator has %d modes, bases have %d and %d", op->nf, rows->nf,
                   cols->nf);
  return QS_OK;
}

static qs_status launch_mf(qs_ctx *ctx, const qs_op *op, const qs_basis *rows, const qs_basis *cols,
                           bool same, const double *x, double *y) {
  if (rows->n == 0) return QS_OK;
  Lookup L = make_lookup(cols);
  // lanes per row: enough threads for the device when rows are few
  int g = 1;
  while (g < 64 && rows->n * g < ((int64_t)1 << 19) && g < op->nterms) g *= 2;
  if (g > 1) {
    const unsigned blocks = (unsigned)((rows->n * g + QS_BLOCK - 1) / QS_BLOCK);
    if (op->nw == 1)
      hipLaunchKernelGGL(k_apply_mf_split<1>, dim3(blocks), dim3(QS_BLOCK), 0, ctx->stream, rows->d_words, rows->n,
                         (const TermD<1> *)op->d_terms, op->nterms, L, (int)same, g, (const double2 *)x,
                         (double2 *)y);
    else
      hipLaunchKernelGGL(k_apply_mf_split<2>, dim3(blocks), dim3(QS_BLOCK), 0, ctx->stream, rows->d_words, rows->n,
                         (const TermD<2> *)op->d_terms, op->nterms, L, (int)same, g, (const double2 *)x,
                         (double2 *)y);
    QS_HIP(ctx, hipGetLastError());
    return QS_OK;
  }
  const unsigned blocks = (unsigned)((rows->n + QS_BLOCK - 1) / QS_BLOCK);
  if (op->nw == 1)
    hipLaunchKernelGGL(k_apply_mf<1>, dim3(blocks), dim3(QS_BLOCK), 0, ctx->stream, rows->d_words, rows->n,
                       (const TermD<1> *)op->d_terms, op->nterms, L, (int)same, (const double2 *)x,
                       (double2 *)y);
  else
    hipLaunchKernelGGL(k_apply_mf<2>, dim3(blocks), dim3(QS_BLOCK), 0, ctx->stream, rows->d_words, rows->n,
                       (const TermD<2> *)op->d_terms, op->nterms, L, (int)same, (const double2 *)x,
                       (double2 *)y);
  QS_HIP(ctx, hipGetLastError());
  return QS_OK;
}

// ---------------------------------------------------------------------------
// dense matrix of a binding (OperatorToDeterminantMatrix, Core/MatrixOperator.cpp:1149-1226):
// M[c * nrows + r] = <r|O|c>, column-major. One lane per row, so the terms of
// a row add into its entries in term order without atomics.
// ---------------------------------------------------------------------------
__global__ void k_zero2(int64_t n, double2 *__restrict__ m) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) m[i] = make_double2(0.0, 0.0);
}

template <int NW>
__global__ void __launch_bounds__(QS_BLOCK)
k_dense(const uint64_t *__restrict__ rows, int64_t nrows, const TermD<NW> *__restrict__ terms, int nterms,
        Lookup L, int same, double2 *__restrict__ M) {
  __shared__ TermD<NW> sh[QS_TILE];
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = row < nrows;
  uint64_t d[NW];
#pragma unroll
  for (int w = 0; w < NW; ++w) d[w] = active ? rows[row * NW + w] : 0;
  for (int base = 0; base < nterms; base += QS_TILE) {
    const int cnt = min(QS_TILE, nterms - base);
    __syncthreads();
    stage_terms<NW>(sh, terms, base, cnt);
    __syncthreads();
    if (active) {
      visit_tile<NW>(d, row, sh, cnt, base, same != 0, L, [&](int tj, int64_t col, uint32_t par, bool) {
        const TermD<NW> &t = sh[tj - base];
        double2 &m = M[col * nrows + row];
        m.x += par ? -t.re : t.re;
        m.y += par ? -t.im : t.im;
      });
    }
  }
}

extern "C" qs_status qs_hop_dense(qs_ctx *ctx, const qs_hop *h, double *M) {
  if (!ctx || !h) return QS_EINVAL;
  const int64_t nr = h->rows->n, nc = h->cols->n;
  if (nr == 0 || nc == 0) return QS_OK;
  if (!M) return QS_EINVAL;
  if (nr > ((int64_t)1 << 31) / (nc > 0 ? nc : 1))
    return qs_fail(ctx, QS_EINVAL, "qs_hop_dense: a %lld x %lld matrix is too large", (long long)nr, (long long)nc);
  QS_HIP(ctx, hipSetDevice(ctx->device));
  const int64_t nel = nr * nc;
  hipLaunchKernelGGL(k_zero2, dim3((unsigned)((nel + 255) / 256)), dim3(256), 0, ctx->stream, nel, (double2 *)M);
  QS_HIP(ctx, hipGetLastError());
  Lookup L = make_lookup(h->cols);
  const unsigned blocks = (unsigned)((nr + QS_BLOCK - 1) / QS_BLOCK);
  if (h->op->nw == 1)
    hipLaunchKernelGGL(k_dense<1>, dim3(blocks), dim3(QS_BLOCK), 0, ctx->stream, h->rows->d_words, nr,
                       (const TermD<1> *)h->op->d_terms, h->op->nterms, L, (int)h->same, (double2 *)M);
  else
    hipLaunchKernelGGL(k_dense<2>, dim3(blocks), dim3(QS_BLOCK), 0, ctx->stream, h->rows->d_words, nr,
                       (const TermD<2> *)h->op->d_terms, h->op->nterms, L, (int)h->same, (double2 *)M);
  QS_HIP(ctx, hipGetLastError());
  if (!ctx->async) QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QS_OK;
}

qs_status qs_launch_mf_ext(qs_ctx *ctx, const qs_hop *h, const double *x, double *y) {
  return launch_mf(ctx, h->op, h->rows, h->cols, h->same, x, y);
}

HopIndex::~HopIndex() {
  for (void *b : buf)
    qs_free(b);
}

static void free_stream(qs_hop *h) {
  if (h->index) {
    h->index.reset();  // the arrays go with the last binding that shares them
  } else {
    void *bufs[] = {h->d_blk_ptr, h->d_blk, h->d_slice_ptr, h->d_slice_len, h->d_rowcnt, h->d_col, h->d_tid};
    for (void *b : bufs)
      qs_free(b);
  }
  qs_free(h->d_diag);
  h->d_inv = nullptr;
  h->d_blk_ptr = nullptr;
  h->d_blk = nullptr;
  h->d_slice_ptr = nullptr;
  h->d_slice_len = nullptr;
  h->d_rowcnt = nullptr;
  h->d_col = nullptr;
  h->d_tid = nullptr;
  h->d_diag = nullptr;
  h->view = SellD();
}

static uint32_t bits_for(uint64_t maxval) {
  uint32_t b = 1;
  while (b < 64 && (maxval >> b) != 0) ++b;
  return b;
}

template <int NW, bool FILL>
static void launch_build(qs_ctx *ctx, qs_hop *h, unsigned blocks, const Lookup &L, unsigned long long *ctr,
                         uint32_t *blk_cnt) {
  const qs_op *op = h->op;
  hipLaunchKernelGGL((k_sell_build<NW, FILL>), dim3(blocks), dim3(QS_BLOCK), 0, ctx->stream, h->rows->d_words,
                     h->rows->n, (const TermD<NW> *)op->d_terms, op->nterms, L, (int)h->same, h->shift_min,
                     h->d_rowcnt, h->d_slice_len, blk_cnt, (void *)h->d_diag, h->view.real_diag, ctr, h->d_blk_ptr,
                     h->d_blk,
                     h->d_slice_ptr, h->d_col, h->d_tid, h->view.packed, h->view.colbits);
}

static qs_status assemble(qs_ctx *ctx, qs_hop *h, size_t budget) {
  const qs_op *op = h->op;
  const qs_basis *rows = h->rows;
  const int64_t n = rows->n;
  const int64_t nslices = (n + 63) / 64;
  h->nslices = nslices;
  h->shift_min = ctx->shift_min;
  if (op->nterms > 65535) return qs_fail(ctx, QS_EUNSUPPORTED, "more than 65535 terms: not assembled");
  if (h->cols->n >= ((int64_t)1 << 31)) return qs_fail(ctx, QS_EUNSUPPORTED, "2^31 or more columns: not assembled");
  QS_TRY(qs_malloc(ctx, (void **)&h->d_slice_ptr, (size_t)(nslices + 1) * 8));
  QS_TRY(qs_malloc(ctx, (void **)&h->d_blk_ptr, (size_t)(nslices + 1) * 8));
  QS_TRY(qs_malloc(ctx, (void **)&h->d_slice_len, (size_t)std::max<int64_t>(nslices, 1) * 4));
  QS_TRY(qs_malloc(ctx, (void **)&h->d_rowcnt, (size_t)std::max<int64_t>(n, 1) * 2));
  // the diagonal is real when every diagonal term is (Hermitian operators)
  bool real_diag = h->same;
  for (const TermHost &t : op->h_terms)
    if ((t.F[0] | t.F[1]) == 0 && t.im != 0.0) real_diag = false;
  h->view.real_diag = real_diag ? 1 : 0;
  const size_t diag_elem = real_diag ? 8 : 16;
  if (h->same) QS_TRY(qs_malloc(ctx, (void **)&h->d_diag, (size_t)std::max<int64_t>(n, 1) * diag_elem));
  // scratch: 3 counters, then the per-slice block counts
  void *scr;
  QS_TRY(qs_scratch(ctx, 64 + (size_t)std::max<int64_t>(nslices, 1) * 4, &scr));
  unsigned long long *ctr = (unsigned long long *)scr;
  uint32_t *blk_cnt = (uint32_t *)((char *)scr + 64);
  QS_HIP(ctx, hipMemsetAsync(ctr, 0, 64, ctx->stream));
  Lookup L = make_lookup(h->cols);
  const unsigned blocks = (unsigned)((n + QS_BLOCK - 1) / QS_BLOCK);
  // the ELL entry format is fixed before the count pass (it does not change the counts)
  const uint32_t colbits = bits_for((uint64_t)std::max<int64_t>(h->cols->n - 1, 1));
  const uint32_t termbits = bits_for((uint64_t)std::max(op->nterms - 1, 1));
  h->view.packed = ctx->ell_pack && colbits + termbits <= 32;
  h->view.colbits = colbits;
  if (n > 0) {
    if (op->nw == 1) launch_build<1, false>(ctx, h, blocks, L, ctr, blk_cnt);
    else launch_build<2, false>(ctx, h, blocks, L, ctr, blk_cnt);
    QS_HIP(ctx, hipGetLastError());
  }
  std::vector<uint32_t> slen((size_t)nslices), bcnt((size_t)nslices);
  unsigned long long c[3] = {0, 0, 0};
  if (nslices) {
    QS_TRY(qs_d2h(ctx, slen.data(), h->d_slice_len, (size_t)nslices * 4));
    QS_TRY(qs_d2h(ctx, bcnt.data(), blk_cnt, (size_t)nslices * 4));
  }
  QS_TRY(qs_d2h(ctx, c, ctr, 24));
  QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<uint64_t> sptr((size_t)nslices + 1, 0), bptr((size_t)nslices + 1, 0);
  for (int64_t s = 0; s < nslices; ++s) {
    sptr[s + 1] = sptr[s] + (uint64_t)slen[s] * 64;
    bptr[s + 1] = bptr[s] + bcnt[s];
  }
  const uint64_t ent = sptr[nslices], nblk = bptr[nslices];
  h->nnz_eff = (int64_t)c[0];
  h->nnz_offdiag = (int64_t)c[1];
  h->block_hits = (int64_t)c[2];
  h->nblocks = (int64_t)nblk;
  h->ell_slots = (int64_t)ent;
  h->nnz_stored = (int64_t)(ent + nblk);
  const uint64_t per_entry = h->view.packed ? 4 : 6;
  h->stream_bytes = (int64_t)(ent * per_entry + nblk * 16 + (uint64_t)n * (2 + 8 * op->nw) +
                              (uint64_t)(nslices + 1) * 20 + (h->same ? (uint64_t)n * diag_elem : 0));
  if (budget && (size_t)h->stream_bytes > budget)
    return qs_fail(ctx, QS_ENOMEM_DEVICE, "assembled stream needs %lld bytes, budget %zu",
                   (long long)h->stream_bytes, budget);
  QS_TRY(qs_malloc(ctx, (void **)&h->d_col, (size_t)std::max<uint64_t>(ent, 1) * 4));
  if (!h->view.packed) QS_TRY(qs_malloc(ctx, (void **)&h->d_tid, (size_t)std::max<uint64_t>(ent, 1) * 2));
  QS_TRY(qs_malloc(ctx, (void **)&h->d_blk, (size_t)std::max<uint64_t>(nblk, 1) * 16));
  QS_TRY(qs_h2d(ctx, h->d_slice_ptr, sptr.data(), sptr.size() * 8));
  QS_TRY(qs_h2d(ctx, h->d_blk_ptr, bptr.data(), bptr.size() * 8));
  if (n > 0) {
    if (op->nw == 1) launch_build<1, true>(ctx, h, blocks, L, ctr, blk_cnt);
    else launch_build<2, true>(ctx, h, blocks, L, ctr, blk_cnt);
    QS_HIP(ctx, hipGetLastError());
    if (ctx->ell_order == 1) {
      if (h->view.packed)
        hipLaunchKernelGGL(k_sell_sort_rows<true>, dim3(blocks), dim3(QS_BLOCK), 0, ctx->stream, n, h->d_rowcnt,
                           h->d_slice_ptr, h->view.colbits, h->d_col, h->d_tid);
      else
        hipLaunchKernelGGL(k_sell_sort_rows<false>, dim3(blocks), dim3(QS_BLOCK), 0, ctx->stream, n, h->d_rowcnt,
                           h->d_slice_ptr, h->view.colbits, h->d_col, h->d_tid);
      QS_HIP(ctx, hipGetLastError());
    }
  }
  QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  SellD &A = h->view;
  A.nrows = n;
  A.nslices = nslices;
  A.nw = op->nw;
  A.nterms = op->nterms;
  A.rows = rows->d_words;
  A.blk_ptr = h->d_blk_ptr;
  A.blk = h->d_blk;
  A.slice_ptr = h->d_slice_ptr;
  A.slice_len = h->d_slice_len;
  A.rowcnt = h->d_rowcnt;
  A.col = h->d_col;
  A.tid = h->d_tid;
  A.val = (const double2 *)op->d_val;
  A.par = op->d_par;
  A.diag = h->d_diag;
  h->index = std::make_shared<HopIndex>();
  h->index->buf = {h->d_blk_ptr, h->d_blk, h->d_slice_ptr, h->d_slice_len, h->d_rowcnt, h->d_col, h->d_tid};
  h->assembled = 1;
  return QS_OK;
}

// Pre-summed diagonal of the rows: for every row, the sum over the diagonal
// terms (listed in `which`, in term order) that hit it of (-1)^popcount(d & P_t) V_t.
// The same terms in the same order as k_sell_build's count pass, so the two
// give bit-identical diagonals.
template <int NW>
__global__ void __launch_bounds__(QS_BLOCK)
k_diag_build(const uint64_t *__restrict__ rows, int64_t nrows, const TermD<NW> *__restrict__ terms,
             const int *__restrict__ which, int nwhich, void *__restrict__ diag, int real_diag) {
  __shared__ TermD<NW> sh[QS_TILE];
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = row < nrows;
  uint64_t d[NW];
#pragma unroll
  for (int w = 0; w < NW; ++w) d[w] = active ? rows[row * NW + w] : 0;
  double dr = 0.0, di = 0.0;
  for (int base = 0; base < nwhich; base += QS_TILE) {
    const int cnt = min(QS_TILE, nwhich - base);
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) sh[i] = terms[which[base + i]];
    __syncthreads();
    for (int j = 0; j < cnt; ++j) {
      const TermD<NW> &t = sh[j];
      bool hit = active;
#pragma unroll
      for (int w = 0; w < NW; ++w) hit = hit && ((d[w] & t.C[w]) == t.V[w]);
      if (hit) {
        uint32_t p = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) p += __popcll(d[w] & t.P[w]);
        dr += (p & 1u) ? -t.re : t.re;
        di += (p & 1u) ? -t.im : t.im;
      }
    }
  }
  if (active) {
    if (real_diag) static_cast<double *>(diag)[row] = dr;
    else static_cast<double2 *>(diag)[row] = make_double2(dr, di);
  }
}

// Pre-summed diagonal of n rows (determinants at `rows`, op->nw words each)
// over the terms listed in d_which (device, nwhich term indices, in term
// order), into diag (doubles when real_diag, else complex).
qs_status qs_diag_rows(qs_ctx *ctx, const qs_op *op, const uint64_t *rows, int64_t n, const int *d_which, int nwhich,
                       void *diag, int real_diag) {
  if (n == 0) return QS_OK;
  const unsigned blocks = (unsigned)((n + QS_BLOCK - 1) / QS_BLOCK);
  if (op->nw == 1)
    hipLaunchKernelGGL(k_diag_build<1>, dim3(blocks), dim3(QS_BLOCK), 0, ctx->stream, rows, n,
                       (const TermD<1> *)op->d_terms, d_which, nwhich, diag, real_diag);
  else
    hipLaunchKernelGGL(k_diag_build<2>, dim3(blocks), dim3(QS_BLOCK), 0, ctx->stream, rows, n,
                       (const TermD<2> *)op->d_terms, d_which, nwhich, diag, real_diag);
  QS_HIP(ctx, hipGetLastError());
  return QS_OK;
}

extern "C" qs_status qs_hop_create_like(qs_ctx *ctx, const qs_op *op, const qs_hop *like, qs_hop **out) {
  if (!ctx || !op || !like || !out) return QS_EINVAL;
  *out = nullptr;
  if (!like->assembled || !like->index)
    return qs_fail(ctx, QS_EINVAL, "qs_hop_create_like needs an assembled binding to share");
  const qs_op *lo = like->op;
  if (op->nf != lo->nf || op->nw != lo->nw)
    return qs_fail(ctx, QS_EINVAL, "qs_hop_create_like: %d modes against the shared binding's %d", op->nf, lo->nf);
  if (op->nterms != lo->nterms)
    return qs_fail(ctx, QS_EINVAL, "qs_hop_create_like: %d terms against the shared binding's %d", op->nterms,
                   lo->nterms);
  for (int t = 0; t < op->nterms; ++t) {
    const TermHost &a = op->h_terms[t], &b = lo->h_terms[t];
    for (int w = 0; w < op->nw; ++w)
      if (a.C[w] != b.C[w] || a.V[w] != b.V[w] || a.F[w] != b.F[w] || a.P[w] != b.P[w])
        return qs_fail(ctx, QS_EINVAL, "qs_hop_create_like: term %d acts differently from the shared binding's", t);
  }
  QS_HIP(ctx, hipSetDevice(ctx->device));
  if (like->view.kind == 1) return qs_prod_create_like(ctx, op, like, out);
  qs_hop *h = new qs_hop();
  h->op = op;
  h->rows = like->rows;
  h->cols = like->cols;
  h->same = like->same;
  h->assembled = 1;
  h->nnz_eff = like->nnz_eff;
  h->nnz_stored = like->nnz_stored;
  h->nnz_offdiag = like->nnz_offdiag;
  h->nblocks = like->nblocks;
  h->block_hits = like->block_hits;
  h->ell_slots = like->ell_slots;
  h->shift_min = like->shift_min;
  h->nslices = like->nslices;
  h->index = like->index;
  h->d_blk_ptr = like->d_blk_ptr;
  h->d_blk = like->d_blk;
  h->d_slice_ptr = like->d_slice_ptr;
  h->d_slice_len = like->d_slice_len;
  h->d_rowcnt = like->d_rowcnt;
  h->d_col = like->d_col;
  h->d_tid = like->d_tid;
  h->view = like->view;
  h->view.val = (const double2 *)op->d_val;
  h->view.par = op->d_par;
  h->view.diag = nullptr;
  const int64_t n = h->rows->n;
  size_t diag_elem = 0;
  qs_status st = QS_OK;
  if (h->same) {
    bool real_diag = true;
    std::vector<int> which;
    for (int t = 0; t < op->nterms; ++t) {
      const TermHost &x = op->h_terms[t];
      if ((x.F[0] | x.F[1]) == 0) {
        which.push_back(t);
        if (x.im != 0.0) real_diag = false;
      }
    }
    h->view.real_diag = real_diag ? 1 : 0;
    diag_elem = real_diag ? 8 : 16;
    st = qs_malloc(ctx, (void **)&h->d_diag, (size_t)std::max<int64_t>(n, 1) * diag_elem);
    void *scr = nullptr;
    if (st == QS_OK && !which.empty()) st = qs_scratch(ctx, which.size() * sizeof(int), &scr);
    if (st == QS_OK && !which.empty()) st = qs_h2d(ctx, scr, which.data(), which.size() * sizeof(int));
    if (st == QS_OK && n > 0) {
      const unsigned blocks = (unsigned)((n + QS_BLOCK - 1) / QS_BLOCK);
      if (op->nw == 1)
        hipLaunchKernelGGL(k_diag_build<1>, dim3(blocks), dim3(QS_BLOCK), 0, ctx->stream, h->rows->d_words, n,
                           (const TermD<1> *)op->d_terms, (const int *)scr, (int)which.size(), (void *)h->d_diag,
                           h->view.real_diag);
      else
        hipLaunchKernelGGL(k_diag_build<2>, dim3(blocks), dim3(QS_BLOCK), 0, ctx->stream, h->rows->d_words, n,
                           (const TermD<2> *)op->d_terms, (const int *)scr, (int)which.size(), (void *)h->d_diag,
                           h->view.real_diag);
      const hipError_t e = hipGetLastError();
      if (e == hipSuccess) {
        const hipError_t e2 = hipStreamSynchronize(ctx->stream);
        if (e2 != hipSuccess) st = qs_hip_fail(ctx, e2, "qs_hop_create_like", __FILE__, __LINE__);
      } else {
        st = qs_hip_fail(ctx, e, "qs_hop_create_like", __FILE__, __LINE__);
      }
    }
    h->view.diag = h->d_diag;
  }
  if (st != QS_OK) {
    free_stream(h);
    delete h;
    return st;
  }
  const uint64_t per_entry = h->view.packed ? 4 : 6;
  h->stream_bytes = (int64_t)((uint64_t)h->ell_slots * per_entry + (uint64_t)h->nblocks * 16 +
                              (uint64_t)n * (2 + 8 * op->nw) + (uint64_t)(h->nslices + 1) * 20 +
                              (uint64_t)n * diag_elem);
  *out = h;
  return QS_OK;
}

extern "C" qs_status qs_hop_create(qs_ctx *ctx, const qs_op *op, const qs_basis *rows, const qs_basis *cols,
                                   int mode, size_t budget, qs_hop **out) {
  if (!ctx || !out) return QS_EINVAL;
  *out = nullptr;
  QS_TRY(check_pair(ctx, op, rows, cols));
  QS_HIP(ctx, hipSetDevice(ctx->device));
  qs_hop *h = new qs_hop();
  h->op = op;
  h->rows = rows;
  h->cols = cols;
  h->same = rows == cols;
  if (mode == QS_HOP_ASSEMBLED || mode == QS_HOP_AUTO) {
    size_t bud = budget;
    if (mode == QS_HOP_AUTO && bud == 0) {
      size_t fr = 0, tot = 0;
      if (hipMemGetInfo(&fr, &tot) == hipSuccess) bud = fr / 2;
    }
    qs_status st = assemble(ctx, h, mode == QS_HOP_ASSEMBLED ? budget : bud);
    if (st != QS_OK) {
      free_stream(h);
      h->assembled = 0;
      if (mode == QS_HOP_ASSEMBLED) {
        delete h;
        return st;
      }
      (void)hipGetLastError();
      ctx->err.clear();
    }
  }
  if (!h->assembled) {
    // nnz_eff for the matrix-free binding is computed lazily by qs_hop_stats
    h->nslices = (rows->n + 63) / 64;
  }
  *out = h;
  return QS_OK;
}

extern "C" qs_status qs_hop_stats(const qs_hop *h, int64_t *nnz_eff, int64_t *nnz_stored,
                                  int64_t *stream_bytes, int *assembled, int64_t *nnz_offdiag) {
  if (!h) return QS_EINVAL;
  if (nnz_offdiag) *nnz_offdiag = h->nnz_offdiag;
  if (nnz_eff) *nnz_eff = h->nnz_eff;
  if (nnz_stored) *nnz_stored = h->nnz_stored;
  if (stream_bytes) *stream_bytes = h->stream_bytes;
  if (assembled) *assembled = h->assembled;
  return QS_OK;
}

__global__ void k_diag_out(int64_t n, const void *__restrict__ diag, int real_diag, double2 *__restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = real_diag ? make_double2(static_cast<const double *>(diag)[i], 0.0)
                       : static_cast<const double2 *>(diag)[i];
}

extern "C" qs_status qs_hop_diagonal(qs_ctx *ctx, const qs_hop *h, double *d) {
  if (!ctx || !h || !d) return QS_EINVAL;
  if (!h->assembled || !h->same || !h->d_diag)
    return qs_fail(ctx, QS_EINVAL, "qs_hop_diagonal needs an assembled binding with rows == cols");
  QS_HIP(ctx, hipSetDevice(ctx->device));
  const int64_t n = h->rows->n;
  if (n > 0) {
    const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + QS_BLOCK - 1) / QS_BLOCK, 4096));
    double2 *out = (double2 *)d;
    if (h->view.kind == 1) QS_TRY(qs_scratch(ctx, (size_t)n * 16, (void **)&out));  // row order first
    hipLaunchKernelGGL(k_diag_out, dim3(blocks), dim3(QS_BLOCK), 0, ctx->stream, n, (const void *)h->d_diag,
                       h->view.real_diag, out);
    QS_HIP(ctx, hipGetLastError());
    if (h->view.kind == 1) QS_TRY(qs_prod_from_rows(ctx, h, 1, out, n, (double2 *)d, n));
  }
  if (!ctx->async) QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QS_OK;
}

extern "C" qs_status qs_hop_layout(const qs_hop *h, int64_t *nblocks, int64_t *block_hits, int64_t *ell_slots,
                                   int *packed, int *shift_min) {
  if (!h) return QS_EINVAL;
  if (nblocks) *nblocks = h->nblocks;
  if (block_hits) *block_hits = h->block_hits;
  if (ell_slots) *ell_slots = h->ell_slots;
  if (packed) *packed = h->view.packed;
  if (shift_min) *shift_min = h->shift_min;
  return QS_OK;
}

// Launch k_spmv for KC interleaved vectors with the term tables in LDS when
// they fit, on no more workgroups than stay resident (grid stride over slices).
template <int KC, int NW, bool PACKED>
static void launch_spmv(qs_ctx *ctx, const qs_hop *h, const double2 *x, int64_t ldx, int64_t ldvx, double2 *y,
                        int64_t ldy, int64_t ldvy) {
  const SellD &A = h->view;
  const size_t shm = term_table_bytes(A.nterms, NW);
  const int64_t want = (A.nslices + 3) / 4;
  if (shm <= (size_t)ctx->term_lds) {
    const void *kern = (const void *)k_spmv<KC, NW, PACKED, true>;
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(want, qs_resident_blocks(ctx, kern, shm)));
    hipLaunchKernelGGL((k_spmv<KC, NW, PACKED, true>), dim3(blocks), dim3(QS_BLOCK), shm, ctx->stream, A, x, ldx,
                       ldvx, y, ldy, ldvy);
  } else {
    const void *kern = (const void *)k_spmv<KC, NW, PACKED, false>;
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(want, qs_resident_blocks(ctx, kern, 0)));
    hipLaunchKernelGGL((k_spmv<KC, NW, PACKED, false>), dim3(blocks), dim3(QS_BLOCK), 0, ctx->stream, A, x, ldx,
                       ldvx, y, ldy, ldvy);
  }
}

template <int KC>
static void launch_spmv_kc(qs_ctx *ctx, const qs_hop *h, const double2 *x, int64_t ldx, int64_t ldvx, double2 *y,
                           int64_t ldy, int64_t ldvy) {
  const SellD &A = h->view;
  if (A.nw == 1) {
    if (A.packed) launch_spmv<KC, 1, true>(ctx, h, x, ldx, ldvx, y, ldy, ldvy);
    else launch_spmv<KC, 1, false>(ctx, h, x, ldx, ldvx, y, ldy, ldvy);
  } else {
    if (A.packed) launch_spmv<KC, 2, true>(ctx, h, x, ldx, ldvx, y, ldy, ldvy);
    else launch_spmv<KC, 2, false>(ctx, h, x, ldx, ldvx, y, ldy, ldvy);
  }
}

// k vectors, entry j of vector v at x[j*ldx + v*ldvx] (y likewise), in
// launches of 4, 2 or 1 vectors
qs_status qs_hop_apply_vectors(qs_ctx *ctx, const qs_hop *h, int k, const double *x, int64_t ldx, int64_t ldvx,
                               double *y, int64_t ldy, int64_t ldvy) {
  const int64_t n = h->rows->n;
  if (n == 0 || k == 0) return QS_OK;
  if (h->view.kind == 1) {
    // product layout: through the binding's row order
    if (ldx != 1 || ldy != 1)
      return qs_fail(ctx, QS_EUNSUPPORTED, "product layout: vectors must be stored one after another");
    double2 *xp;
    QS_TRY(qs_scratch(ctx, (size_t)2 * k * n * 16, (void **)&xp));
    double2 *yp = xp + (size_t)k * n;
    QS_TRY(qs_prod_to_rows(ctx, h, k, (const double2 *)x, ldvx, xp, n));
    QS_TRY(qs_prod_apply_rows(ctx, h, k, xp, n, yp, n));
    return qs_prod_from_rows(ctx, h, k, yp, n, (double2 *)y, ldvy);
  }
  for (int v0 = 0; v0 < k;) {
    const int kc = k - v0 >= 4 ? 4 : (k - v0 >= 2 ? 2 : 1);
    const double2 *xs = (const double2 *)x + v0 * ldvx;
    double2 *ys = (double2 *)y + v0 * ldvy;
    if (kc == 4) launch_spmv_kc<4>(ctx, h, xs, ldx, ldvx, ys, ldy, ldvy);
    else if (kc == 2) launch_spmv_kc<2>(ctx, h, xs, ldx, ldvx, ys, ldy, ldvy);
    else launch_spmv_kc<1>(ctx, h, xs, ldx, ldvx, ys, ldy, ldvy);
    QS_HIP(ctx, hipGetLastError());
    v0 += kc;
  }
  return QS_OK;
}

// vectors interleaved: entry j of vector v at (j*ld + v)
qs_status qs_hop_apply_strided(qs_ctx *ctx, const qs_hop *h, int k, const double *x, int64_t ldx,
                               double *y, int64_t ldy) {
  return qs_hop_apply_vectors(ctx, h, k, x, ldx, 1, y, ldy, 1);
}

extern "C" qs_status qs_hop_apply(qs_ctx *ctx, const qs_hop *h, const double *x, double *y) {
  if (!ctx || !h || (h->rows->n && (!x || !y))) return QS_EINVAL;
  QS_HIP(ctx, hipSetDevice(ctx->device));
  if (h->assembled) {
    QS_TRY(qs_hop_apply_strided(ctx, h, 1, x, 1, y, 1));
  } else {
    QS_TRY(launch_mf(ctx, h->op, h->rows, h->cols, h->same, x, y));
  }
  if (!ctx->async) QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QS_OK;
}

extern "C" qs_status qs_hop_apply_multi(qs_ctx *ctx, const qs_hop *h, int k, const double *x, double *y) {
  if (!ctx || !h || k < 0) return QS_EINVAL;
  QS_HIP(ctx, hipSetDevice(ctx->device));
  const int64_t nc = h->cols->n, nr = h->rows->n;
  if (h->assembled) {
    QS_TRY(qs_hop_apply_vectors(ctx, h, k, x, 1, nc, y, 1, nr));
  } else {
    for (int v = 0; v < k; ++v) QS_TRY(launch_mf(ctx, h->op, h->rows, h->cols, h->same, x + 2 * v * nc, y + 2 * v * nr));
  }
  if (!ctx->async) QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QS_OK;
}

extern "C" qs_status qs_hop_destroy(qs_hop *h) {
  if (!h) return QS_OK;
  free_stream(h);
  delete h;
  return QS_OK;
}

extern "C" qs_status qs_apply(qs_ctx *ctx, const qs_op *op, const qs_basis *rows, const qs_basis *cols,
                              const double *x, double *y) {
  if (!ctx) return QS_EINVAL;
  QS_TRY(check_pair(ctx, op, rows, cols));
  QS_HIP(ctx, hipSetDevice(ctx->device));
  QS_TRY(launch_mf(ctx, op, rows, cols, rows == cols, x, y));
  if (!ctx->async) QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QS_OK;
}

// ---------------------------------------------------------------------------
// <x|O|x> without any allocation: the product O x and the per-block partial
// sums live in `scratch`, a file-scope device array that every GPU the
// library runs on holds until the process ends. Each call overwrites all of
// it that it reads (O x over every row, one partial per block), so nothing has
// to start at zero; calls that overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------
#define QS_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
#define QS_NOALLOC_BLOCKS 1024
static __device__ __align__(256) unsigned char scratch[QS_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_addr;  // device -> address of `scratch` on that device

// The device address of `scratch` on ctx's device, looked up once per device
// with that device current (each GPU holds its own copy at its own address).
qs_status scratch_address(qs_ctx *ctx, void **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_addr.find(ctx->device);
  if (it == g_scratch_addr.end()) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return qs_fail(ctx, QS_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", ctx->device,
                     hipGetErrorString(e));
    it = g_scratch_addr.emplace(ctx->device, p).first;
  }
  *out = it->second;
  return QS_OK;
}
}  // namespace

// partial[b] = sum over block b's rows of conj(x_i) y_i (re at [b], im at [nb + b])
__global__ void __launch_bounds__(QS_BLOCK)
k_expect_partial(int64_t n, const double2 *__restrict__ x, const double2 *__restrict__ y, double *__restrict__ part) {
  __shared__ double red[QS_BLOCK / 64];
  double sr = 0.0, si = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 a = x[i], c = y[i];
    sr += a.x * c.x + a.y * c.y;
    si += a.x * c.y - a.y * c.x;
  }
  const double r = block_sum_values(sr, red);
  const double m = block_sum_values(si, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = r;
    part[gridDim.x + blockIdx.x] = m;
  }
}

// out[0] + i out[1] = the fixed-order sum of the nb partials
__global__ void __launch_bounds__(QS_BLOCK)
k_expect_finish(int nb, const double *__restrict__ part, double *__restrict__ out) {
  __shared__ double red[QS_BLOCK / 64];
  const double r = block_sum_fixed(part, nb, red);
  const double m = block_sum_fixed(part + nb, nb, red);
  if (threadIdx.x == 0) {
    out[0] = r;
    out[1] = m;
  }
}

extern "C" qs_status qs_expect_noalloc(qs_ctx *ctx, const qs_op *op, const qs_basis *b, const double *x,
                                       double *out) {
  if (!ctx) return QS_EINVAL;
  if (!x || !out) return qs_fail(ctx, QS_EINVAL, "qs_expect_noalloc: null vector or output");
  QS_TRY(check_pair(ctx, op, b, b));
  const int64_t n = b->n;
  const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(QS_NOALLOC_BLOCKS, (n + QS_BLOCK - 1) / QS_BLOCK));
  const size_t vec_bytes = ((size_t)n * 16 + 255) / 256 * 256;
  const size_t need = vec_bytes + (size_t)nb * 16;
  if (need > QS_NOALLOC_SCRATCH_BYTES)
    return qs_fail(ctx, QS_EINVAL,
                   "qs_expect_noalloc: a basis of %lld determinants needs %zu bytes of scratch; the static "
                   "scratch holds %zu bytes (at most %lld determinants); use qs_expect",
                   (long long)n, need, QS_NOALLOC_SCRATCH_BYTES,
                   (long long)((QS_NOALLOC_SCRATCH_BYTES - QS_NOALLOC_BLOCKS * 16) / 16));
  QS_HIP(ctx, hipSetDevice(ctx->device));
  void *base = nullptr;
  QS_TRY(scratch_address(ctx, &base));
  double *y = (double *)base;
  double *part = (double *)((char *)base + vec_bytes);
  QS_TRY(launch_mf(ctx, op, b, b, true, x, y));
  hipLaunchKernelGGL(k_expect_partial, dim3(nb), dim3(QS_BLOCK), 0, ctx->stream, n, (const double2 *)x,
                     (const double2 *)y, part);
  hipLaunchKernelGGL(k_expect_finish, dim3(1), dim3(QS_BLOCK), 0, ctx->stream, nb, (const double *)part, out);
  QS_HIP(ctx, hipGetLastError());
  if (!ctx->async) QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QS_OK;
}

extern "C" qs_status qs_expect(qs_ctx *ctx, const qs_op *op, const qs_basis *b, const double *x, double *re,
                               double *im) {
  if (!ctx || !re || !im) return QS_EINVAL;
  QS_TRY(check_pair(ctx, op, b, b));
  QS_HIP(ctx, hipSetDevice(ctx->device));
  double *y = nullptr;
  QS_TRY(qs_malloc(ctx, (void **)&y, (size_t)std::max<int64_t>(b->n, 1) * 16));
  qs_status st = launch_mf(ctx, op, b, b, true, x, y);
  // <x|O x>
  if (st == QS_OK) st = qs_dot_internal(ctx, b->n, x, y, re, im);
  qs_free(y);
  return st;
}
