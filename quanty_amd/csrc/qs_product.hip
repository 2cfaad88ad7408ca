// Product layout of an operator on a restricted sector.
//
// The modes are split into two groups, A (e.g. an impurity shell) and B (its
// bath shells). When every restriction of the sector lies inside A, inside B,
// or covers all modes, the sector is a union of products: each allowed
// configuration b of the B modes pairs with every allowed configuration a of
// the A modes whose count fits the total window (its "A-set", which depends
// on b's count only). Rows are ordered by (b, a): a group of consecutive rows
// per b, one row per a of its A-set.
//
// Every ladder string t acts as (A part) x (B part), and its sign
// (-1)^popcount(d & P_t) factorizes into popcount(a & P_A) + popcount(b & P_B).
// The pull form y = O x then splits into
//   * terms that leave B alone (C_B = P_B = 0): per A-set a small matrix over
//     the A-set's configurations (the "A-matrix", values and signs folded in),
//     applied inside the row's own group;
//   * terms that move or test B: stored once per group as an entry (source
//     group, term, sign from b); a lane applies it to its own a, unchanged
//     (identity map: consecutive lanes read consecutive source rows, one
//     coalesced load per entry) or through a map a -> source index.
// On an impurity with bath shells the bath moves are most of the entries, so
// most of the gathers become contiguous row reads shared by a wavefront; the
// index stream shrinks to a few bytes per row. The diagonal stays per row.
//
// The result equals OperatorPsiConserveBasis (Core/OperatorWaveFunction.cpp:245)
// on the sector, as the assembled SELL layout does; only the row order differs:
// vectors are converted at the binding's boundary (qs_hop_apply and the
// Lanczos entry points take and return canonical order).
#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "qs_internal.h"
#include "qs_kernels.h"

namespace {

// mode i of a determinant: word i >> 6, bit 63 - (i & 63)
inline bool det_bit(const uint64_t *w, int i) { return (w[i >> 6] >> (63 - (i & 63))) & 1ULL; }

// Values over a mode subset: modes[0] at the highest of nm bits, so numeric
// order over a subset is the canonical order restricted to it.
struct Subset {
  int nw = 1;
  std::vector<int> modes;
  // deposit tables: byte j of a value (bits 8j..8j+7) -> determinant words
  std::vector<uint64_t> dep;  // 8 * 256 * nw
  int nm() const { return (int)modes.size(); }
  void init(int nw_, const std::vector<int> &m) {
    nw = nw_;
    modes = m;
    dep.assign((size_t)8 * 256 * nw, 0);
    const int n = nm();
    for (int j = 0; j < 8; ++j)
      for (int v = 0; v < 256; ++v)
        for (int k0 = 0; k0 < 8; ++k0) {
          const int k = 8 * j + k0;
          if (k < n && ((v >> k0) & 1)) {
            const int mode = modes[n - 1 - k];
            dep[((size_t)j * 256 + v) * nw + (mode >> 6)] |= 1ULL << (63 - (mode & 63));
          }
        }
  }
  uint64_t extract(const uint64_t *w) const {
    uint64_t v = 0;
    for (int mode : modes) v = (v << 1) | (det_bit(w, mode) ? 1ULL : 0ULL);
    return v;
  }
  void deposit(uint64_t v, uint64_t *w) const {
    for (int j = 0; j < 8 && v; ++j, v >>= 8) {
      const uint64_t *e = &dep[((size_t)j * 256 + (v & 0xFF)) * nw];
      for (int k = 0; k < nw; ++k) w[k] |= e[k];
    }
  }
};

struct Window {
  uint64_t mask;  // over a subset
  int lo, hi;
};

// Values over nm modes satisfying every window, ascending. False when more
// than `limit` would be produced.
bool enumerate_subset(int nm, const std::vector<Window> &ws, std::vector<uint64_t> &out, size_t limit) {
  const int nr = (int)ws.size();
  std::vector<int> cnt(nr, 0);
  // rem[p * nr + r]: modes of window r at positions >= p (position p = bit nm-1-p)
  std::vector<int> rem((size_t)(nm + 1) * std::max(nr, 1), 0);
  for (int p = nm - 1; p >= 0; --p)
    for (int r = 0; r < nr; ++r)
      rem[(size_t)p * nr + r] = rem[(size_t)(p + 1) * nr + r] + (int)((ws[r].mask >> (nm - 1 - p)) & 1ULL);
  auto viable = [&](int p) {
    for (int r = 0; r < nr; ++r)
      if (cnt[r] > ws[r].hi || cnt[r] + rem[(size_t)p * nr + r] < ws[r].lo) return false;
    return true;
  };
  if (!viable(0)) return true;
  // iterative depth-first search: choice[p] = bit chosen at position p (0 first)
  std::vector<int> choice(nm + 1, -1);
  uint64_t v = 0;
  int p = 0;
  if (nm == 0) {
    out.push_back(0);
    return true;
  }
  while (p >= 0) {
    if (choice[p] == 1) {  // both branches done: undo and go up
      const uint64_t bit = 1ULL << (nm - 1 - p);
      for (int r = 0; r < nr; ++r)
        if (ws[r].mask & bit) --cnt[r];
      v &= ~bit;
      choice[p] = -1;
      --p;
      continue;
    }
    const uint64_t bit = 1ULL << (nm - 1 - p);
    if (choice[p] == 0) {  // take bit 1 next
      choice[p] = 1;
      v |= bit;
      for (int r = 0; r < nr; ++r)
        if (ws[r].mask & bit) ++cnt[r];
    } else {
      choice[p] = 0;
    }
    if (!viable(p + 1)) continue;
    if (p + 1 == nm) {
      out.push_back(v);
      if (out.size() > limit) return false;
      continue;
    }
    ++p;
  }
  return true;
}

inline int parity64(uint64_t x) { return __builtin_popcountll(x) & 1; }

// index of v in the ascending list s, or -1
inline int32_t find_sorted(const std::vector<uint64_t> &s, uint64_t v) {
  auto it = std::lower_bound(s.begin(), s.end(), v);
  return (it != s.end() && *it == v) ? (int32_t)(it - s.begin()) : -1;
}

struct TermParts {
  int t;
  uint64_t CA, VA, FA, PA, CB, VB, FB, PB;
};

struct ProdBuild {
  // per A-set
  std::vector<std::vector<uint64_t>> aset;
  // per group
  std::vector<uint64_t> gb;
  std::vector<int> gset;
  std::vector<int64_t> gpos;
};

}  // namespace

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// y = O x in the binding's row order for KC vectors (entry j of vector k at
// x[j*ldx + k*ldvx]), one wavefront per slice, XCD-aware slice ranges.
template <int KC, bool LDS>
__global__ void __launch_bounds__(QS_BLOCK)
k_prod(SellD A, const double2 *__restrict__ x, int64_t ldx, int64_t ldvx, double2 *__restrict__ y, int64_t ldy,
       int64_t ldvy) {
  extern __shared__ char smem[];
  const double2 *tv = A.val;
  if (LDS) {
    stage_term_values(A, smem);
    __syncthreads();
    tv = reinterpret_cast<const double2 *>(smem);
  }
  double2 *xs = prod_stage_area(smem, A.nterms, LDS, KC);
  const int lane = threadIdx.x & 63;
  int64_t s0, s1, ds;
  xcd_items(A.nslices, s0, s1, ds);
  s0 = wave_uniform(s0);
  for (int64_t s = s0; s < s1; s += ds) {
    double ar[KC], ai[KC];
    const int64_t r = prod_row<KC>(A, s, lane, tv, x, ldx, ldvx, ar, ai, xs);
    if (r >= 0) {
#pragma unroll
      for (int k = 0; k < KC; ++k) y[r * ldy + k * ldvy] = make_double2(ar[k], ai[k]);
    }
  }
}

// canonical -> row order: xp[r + k*ldvp] = x[inv[r] + k*ldvx]
__global__ void k_prod_to_rows(int64_t n, int k, const int32_t *__restrict__ inv, const double2 *__restrict__ x,
                               int64_t ldvx, double2 *__restrict__ xp, int64_t ldvp) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = inv[r];
    for (int v = 0; v < k; ++v) xp[r + v * ldvp] = x[c + v * ldvx];
  }
}

// row order -> canonical: y[inv[r] + k*ldvy] = yp[r + k*ldvp]
__global__ void k_prod_from_rows(int64_t n, int k, const int32_t *__restrict__ inv, const double2 *__restrict__ yp,
                                 int64_t ldvp, double2 *__restrict__ y, int64_t ldvy) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = inv[r];
    for (int v = 0; v < k; ++v) y[c + v * ldvy] = yp[r + v * ldvp];
  }
}

// row order words from the canonical basis: pw[r] = words[inv[r]]
__global__ void k_prod_words(int64_t n, int nw, const int32_t *__restrict__ inv, const uint64_t *__restrict__ words,
                             uint64_t *__restrict__ pw) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
    for (int w = 0; w < nw; ++w) pw[r * nw + w] = words[(int64_t)inv[r] * nw + w];
}

static unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }

// ---------------------------------------------------------------------------
// host: per-binding values (diagonal, A-matrix)
// ---------------------------------------------------------------------------
// The binding's diagonal over its rows, from op's compiled terms and the
// shared index (the entries and A-matrix slots name terms; their values come
// from op's term table at run time).
static qs_status prod_bind_values(qs_ctx *ctx, qs_hop *h) {
  const qs_op *op = h->op;
  const ProdHost &P = *h->index->prod;
  const int64_t n = h->rows->n;
  // diagonal: the diagonal terms summed per row, as the SELL count pass does
  bool real_diag = true;
  for (int t : P.diag_terms)
    if (op->h_terms[t].im != 0.0) real_diag = false;
  const size_t el = real_diag ? 8 : 16;
  QS_TRY(qs_malloc(ctx, (void **)&h->d_diag, (size_t)std::max<int64_t>(n, 1) * el));
  h->view.real_diag = real_diag ? 1 : 0;
  if (n > 0) {
    const int nw = op->nw;
    const size_t wbytes = (size_t)n * nw * 8;
    const size_t tbytes = std::max<size_t>(P.diag_terms.size(), 1) * sizeof(int);
    void *scr;
    QS_TRY(qs_scratch(ctx, wbytes + tbytes, &scr));
    uint64_t *pw = (uint64_t *)scr;
    int *which = (int *)((char *)scr + wbytes);
    hipLaunchKernelGGL(k_prod_words, dim3(grid_for(n)), dim3(256), 0, ctx->stream, n, nw, h->d_inv,
                       h->rows->d_words, pw);
    QS_HIP(ctx, hipGetLastError());
    if (!P.diag_terms.empty()) QS_TRY(qs_h2d(ctx, which, P.diag_terms.data(), P.diag_terms.size() * sizeof(int)));
    QS_TRY(qs_diag_rows(ctx, op, pw, n, which, (int)P.diag_terms.size(), h->d_diag, h->view.real_diag));
    QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  h->view.diag = h->d_diag;
  h->view.val = (const double2 *)op->d_val;
  h->view.par = op->d_par;
  h->view.nterms = op->nterms;
  return QS_OK;
}

// ---------------------------------------------------------------------------
// host: the shared index
// ---------------------------------------------------------------------------
static qs_status prod_build(qs_ctx *ctx, qs_hop *h, const uint8_t *amask) {
  const qs_op *op = h->op;
  const qs_basis *B = h->rows;
  const int nf = B->nf, nw = B->nw;
  const int64_t n = B->n;
  if (!B->has_rank)
    return qs_fail(ctx, QS_EUNSUPPORTED, "product layout: the basis must be a restricted sector (qs_basis_enumerate)");
  if (op->nterms > 32767) return qs_fail(ctx, QS_EUNSUPPORTED, "product layout: more than 32767 terms");
  if (n >= ((int64_t)1 << 31)) return qs_fail(ctx, QS_EUNSUPPORTED, "product layout: 2^31 or more rows");
  // mode groups
  std::vector<int> am, bm;
  std::vector<char> inA(nf, 0);
  for (int i = 0; i < nf; ++i)
    if ((amask[i >> 3] >> (7 - (i & 7))) & 1) inA[i] = 1;
  for (int i = 0; i < nf; ++i) (inA[i] ? am : bm).push_back(i);
  if (am.empty() || bm.empty() || am.size() > 62 || bm.size() > 63)
    return qs_fail(ctx, QS_EINVAL, "product layout: %zu modes in group A and %zu in group B (each 1..62)", am.size(),
                   bm.size());
  Subset SA, SB;
  SA.init(nw, am);
  SB.init(nw, bm);
  // restriction windows
  std::vector<Window> wa, wb;
  int flo = 0, fhi = nf;
  const int nbytes = (nf + 7) / 8;
  for (int r = 0; r < B->nres; ++r) {
    const uint8_t *m = &B->masks[(size_t)r * nbytes];
    int na = 0, nb = 0;
    uint64_t ma = 0, mb = 0;
    for (int k = 0; k < (int)am.size(); ++k)
      if ((m[am[k] >> 3] >> (7 - (am[k] & 7))) & 1) {
        ++na;
        ma |= 1ULL << (am.size() - 1 - k);
      }
    for (int k = 0; k < (int)bm.size(); ++k)
      if ((m[bm[k] >> 3] >> (7 - (bm[k] & 7))) & 1) {
        ++nb;
        mb |= 1ULL << (bm.size() - 1 - k);
      }
    if (na == 0 && nb == 0) continue;
    if (nb == 0) {
      wa.push_back({ma, B->lo[r], B->hi[r]});
    } else if (na == 0) {
      wb.push_back({mb, B->lo[r], B->hi[r]});
    } else if (na == (int)am.size() && nb == (int)bm.size()) {
      flo = std::max(flo, B->lo[r]);
      fhi = std::min(fhi, B->hi[r]);
    } else {
      return qs_fail(ctx, QS_EUNSUPPORTED,
                     "product layout: restriction %d holds modes of both groups without covering all modes", r);
    }
  }
  // the total window bounds each group's own count too
  const int nA = (int)am.size(), nB = (int)bm.size();
  wa.push_back({nA == 64 ? ~0ULL : (1ULL << nA) - 1, std::max(0, flo - nB), fhi});
  wb.push_back({nB == 64 ? ~0ULL : (1ULL << nB) - 1, std::max(0, flo - nA), fhi});
  std::vector<uint64_t> acfg, bcfg;
  if (!enumerate_subset(nA, wa, acfg, (size_t)1 << 22))
    return qs_fail(ctx, QS_EUNSUPPORTED, "product layout: more than 2^22 allowed configurations of group A");
  if (!enumerate_subset(nB, wb, bcfg, (size_t)4 * n + ((size_t)1 << 20)))
    return qs_fail(ctx, QS_EUNSUPPORTED, "product layout: the B configurations far outnumber the sector");
  // A-sets keyed by the count window left for A
  ProdBuild pb;
  std::map<std::pair<int, int>, int> setid;
  std::vector<int> set_of_nb(bm.size() + 1, -1);
  for (int k = 0; k <= (int)bm.size(); ++k) {
    const int lo = std::max(0, flo - k), hi = std::min((int)am.size(), fhi - k);
    if (lo > hi) continue;
    auto key = std::make_pair(lo, hi);
    auto it = setid.find(key);
    if (it == setid.end()) {
      std::vector<uint64_t> s;
      for (uint64_t a : acfg) {
        const int c = __builtin_popcountll(a);
        if (c >= lo && c <= hi) s.push_back(a);
      }
      if (s.empty()) continue;
      if (s.size() >= 65535) return qs_fail(ctx, QS_EUNSUPPORTED, "product layout: an A-set of 65535 or more configurations");
      setid[key] = (int)pb.aset.size();
      pb.aset.push_back(std::move(s));
      it = setid.find(key);
    }
    set_of_nb[k] = it->second;
  }
  if (pb.aset.size() > 65535) return qs_fail(ctx, QS_EUNSUPPORTED, "product layout: too many A-sets");
  int64_t rows = 0;
  for (uint64_t b : bcfg) {
    const int g = set_of_nb[__builtin_popcountll(b)];
    if (g < 0) continue;
    pb.gb.push_back(b);
    pb.gset.push_back(g);
    pb.gpos.push_back(rows);
    rows += (int64_t)pb.aset[g].size();
  }
  if (rows != n)
    return qs_fail(ctx, QS_EINVAL, "product layout: the groups hold %lld rows, the sector %lld", (long long)rows,
                   (long long)n);
  const int64_t ngroups = (int64_t)pb.gb.size();
  std::unordered_map<uint64_t, int32_t> gid;
  gid.reserve((size_t)ngroups * 2);
  for (int64_t k = 0; k < ngroups; ++k) gid[pb.gb[k]] = (int32_t)k;
  // slices
  std::vector<int4> slices;
  for (int64_t k = 0; k < ngroups; ++k) {
    const int m = (int)pb.aset[pb.gset[k]].size();
    for (int c = 0; c * 64 < m; ++c)
      slices.push_back(make_int4((int)(pb.gpos[k] + 64 * c), std::min(64, m - 64 * c), (int)k, pb.gset[k] | (c << 16)));
  }
  // canonical index of every row: the row's determinant looked up in the basis
  std::vector<uint64_t> pw((size_t)n * nw, 0);
  for (int64_t k = 0; k < ngroups; ++k) {
    uint64_t wb_[QS_MAX_WORDS] = {0, 0};
    SB.deposit(pb.gb[k], wb_);
    const std::vector<uint64_t> &as = pb.aset[pb.gset[k]];
    for (size_t i = 0; i < as.size(); ++i) {
      uint64_t *w = &pw[((size_t)pb.gpos[k] + i) * nw];
      for (int j = 0; j < nw; ++j) w[j] = wb_[j];
      SA.deposit(as[i], w);
    }
  }
  std::vector<int32_t> inv((size_t)std::max<int64_t>(n, 1), 0);
  if (n > 0) {
    void *scr;
    QS_TRY(qs_scratch(ctx, (size_t)n * nw * 8 + (size_t)n * 8, &scr));
    uint64_t *dpw = (uint64_t *)scr;
    int64_t *didx = (int64_t *)((char *)scr + (size_t)n * nw * 8);
    QS_TRY(qs_h2d(ctx, dpw, pw.data(), (size_t)n * nw * 8));
    QS_TRY(qs_basis_index_dev(ctx, B, dpw, n, didx));
    std::vector<int64_t> idx((size_t)n);
    QS_TRY(qs_d2h(ctx, idx.data(), didx, (size_t)n * 8));
    std::vector<char> seen((size_t)n, 0);
    for (int64_t r = 0; r < n; ++r) {
      const int64_t c = idx[(size_t)r];
      if (c < 0 || c >= n || seen[(size_t)c])
        return qs_fail(ctx, QS_EINVAL, "product layout: row %lld is not a distinct determinant of the sector",
                       (long long)r);
      seen[(size_t)c] = 1;
      inv[(size_t)r] = (int32_t)c;
    }
  }
  pw.clear();
  pw.shrink_to_fit();
  // terms
  auto ProdH = std::make_shared<ProdHost>();
  ProdHost &P = *ProdH;
  std::vector<TermParts> aterms, bterms;
  for (int t = 0; t < op->nterms; ++t) {
    const TermHost &th = op->h_terms[t];
    bool diag = true;
    for (int w = 0; w < nw; ++w) diag = diag && th.F[w] == 0;
    if (diag) {
      P.diag_terms.push_back(t);
      continue;
    }
    TermParts tp{t, SA.extract(th.C), SA.extract(th.V), SA.extract(th.F), SA.extract(th.P),
                 SB.extract(th.C), SB.extract(th.V), SB.extract(th.F), SB.extract(th.P)};
    if (tp.CB == 0 && tp.PB == 0) aterms.push_back(tp);
    else bterms.push_back(tp);
  }
  int64_t nnz = 0;
  // A-matrices: per A-set and row a, one 4-byte slot per hit of a term on A
  // alone (source index | term << 16 | sign << 31), column-major. Each A-set
  // serves every group that holds it, so its hits count once per such group.
  std::vector<int64_t> set_groups(pb.aset.size(), 0), set_chunks(pb.aset.size(), 0);
  for (int g : pb.gset) ++set_groups[(size_t)g];
  std::vector<int4> asets(pb.aset.size());
  std::vector<uint32_t> hsrc;
  double cost = 0.0;
  for (size_t g = 0; g < pb.aset.size(); ++g) {
    const std::vector<uint64_t> &as = pb.aset[g];
    const int m = (int)as.size();
    set_chunks[g] = (m + 63) / 64;
    std::vector<std::vector<uint32_t>> rowslots((size_t)m);
    for (const TermParts &tp : aterms)
      for (int i = 0; i < m; ++i) {
        const uint64_t a = as[i];
        if ((a & tp.CA) != tp.VA) continue;
        const int32_t j = find_sorted(as, a ^ tp.FA);
        if (j < 0) continue;
        nnz += set_groups[g];
        rowslots[(size_t)i].push_back((uint32_t)j | ((uint32_t)tp.t << 16) | ((uint32_t)parity64(a & tp.PA) << 31));
      }
    int k = 0;
    for (auto &rs : rowslots) k = std::max(k, (int)rs.size());
    const int64_t off = (int64_t)hsrc.size();
    asets[g] = make_int4(m, k, (int)off, 0);
    hsrc.resize((size_t)(off + (int64_t)k * m), 0xFFFFu);
    for (int i = 0; i < m; ++i)
      for (int e = 0; e < (int)rowslots[(size_t)i].size(); ++e)
        hsrc[(size_t)(off + (int64_t)e * m + i)] = rowslots[(size_t)i][(size_t)e];
    // a slot is a 4-byte load (a quarter of a 16-byte one)
    cost += 0.25 * k * set_chunks[g] * set_groups[g];
  }
  if (hsrc.size() >= ((size_t)1 << 31)) return qs_fail(ctx, QS_EUNSUPPORTED, "product layout: A-matrices too large");
  // entries and maps: per group, the B-moving terms with a source in the sector
  std::map<std::tuple<uint64_t, uint64_t, uint64_t, uint64_t, int, int>, int> mapid;
  std::vector<int32_t> amap;
  std::vector<uint32_t> amap_off;
  std::vector<int64_t> map_hits;
  std::vector<uint2> eptr((size_t)ngroups + 1, make_uint2(0u, 0u));
  std::vector<uint4> ent;
  struct Hit {
    uint32_t pos, moff, code;
  };
  std::vector<Hit> gh;
  int64_t nmapped = 0;
  for (int64_t k = 0; k < ngroups; ++k) {
    const uint64_t b = pb.gb[k];
    const int g = pb.gset[k];
    gh.clear();
    for (const TermParts &tp : bterms) {
      if ((b & tp.CB) != tp.VB) continue;
      auto it = gid.find(b ^ tp.FB);
      if (it == gid.end()) continue;
      const int kk = it->second;
      const int gs = pb.gset[kk];
      uint32_t moff;
      int64_t hits;
      if (tp.CA == 0 && tp.PA == 0 && gs == g) {
        moff = 0xFFFFFFFFu;  // identity: every lane reads its own index
        hits = (int64_t)pb.aset[g].size();
      } else {
        auto key = std::make_tuple(tp.CA, tp.VA, tp.FA, tp.PA, g, gs);
        auto mi = mapid.find(key);
        if (mi == mapid.end()) {
          const std::vector<uint64_t> &ad = pb.aset[g], &asrc = pb.aset[gs];
          std::vector<int32_t> mv(ad.size(), -1);
          int64_t h = 0;
          for (size_t i = 0; i < ad.size(); ++i) {
            const uint64_t a = ad[i];
            if ((a & tp.CA) != tp.VA) continue;
            const int32_t j = find_sorted(asrc, a ^ tp.FA);
            if (j < 0) continue;
            mv[i] = j | (parity64(a & tp.PA) << 30);
            ++h;
          }
          const int id = (int)amap_off.size();
          if (amap.size() + mv.size() >= 0xFFFFFFFFull)
            return qs_fail(ctx, QS_EUNSUPPORTED, "product layout: maps too large");
          mapid[key] = id;
          amap_off.push_back((uint32_t)amap.size());
          amap.insert(amap.end(), mv.begin(), mv.end());
          map_hits.push_back(h);
          mi = mapid.find(key);
        }
        hits = map_hits[(size_t)mi->second];
        if (hits == 0) continue;
        moff = amap_off[(size_t)mi->second];
      }
      nnz += hits;
      gh.push_back({(uint32_t)pb.gpos[kk], moff, (uint32_t)tp.t | ((uint32_t)parity64(b & tp.PB) << 31)});
    }
    // identity entries first, then mapped ones; each part in source-row
    // order, so neighbouring entries read neighbouring rows
    std::sort(gh.begin(), gh.end(), [](const Hit &p, const Hit &q) {
      const bool pm = p.moff != 0xFFFFFFFFu, qm = q.moff != 0xFFFFFFFFu;
      if (pm != qm) return qm;
      if (p.pos != q.pos) return p.pos < q.pos;
      if (p.moff != q.moff) return p.moff < q.moff;
      return (p.code & 0x7FFFFFFFu) < (q.code & 0x7FFFFFFFu);
    });
    const uint32_t first = (uint32_t)ent.size();
    uint32_t nid = 0;
    for (const Hit &e : gh) {
      ent.push_back(make_uint4(e.pos, e.code, e.moff, 0u));
      if (e.moff == 0xFFFFFFFFu) ++nid;
      else ++nmapped;
    }
    if (ent.size() >= 0xFFFFFFFFull)
      return qs_fail(ctx, QS_EUNSUPPORTED, "product layout: 2^32 or more entries");
    eptr[(size_t)k] = make_uint2(first, first + nid);
    eptr[(size_t)k + 1] = make_uint2((uint32_t)ent.size(), (uint32_t)ent.size());
    // an identity entry is one 16-byte load per chunk, a mapped one two
    const uint32_t nent = (uint32_t)ent.size() - first;
    cost += (double)set_chunks[(size_t)g] * (nid + 2.0 * (nent - nid) + 2.0);
  }
  // diagonal hits, counted per group from the factorized conditions
  int64_t dhits = 0;
  for (int t : P.diag_terms) {
    const TermHost &th = op->h_terms[t];
    const uint64_t CA = SA.extract(th.C), VA = SA.extract(th.V), CB = SB.extract(th.C), VB = SB.extract(th.V);
    std::vector<int64_t> per_set(pb.aset.size(), 0);
    for (size_t g = 0; g < pb.aset.size(); ++g)
      for (uint64_t a : pb.aset[g]) per_set[g] += (a & CA) == VA;
    for (int64_t k = 0; k < ngroups; ++k)
      if ((pb.gb[k] & CB) == VB) dhits += per_set[(size_t)pb.gset[k]];
  }
  P.ngroups = ngroups;
  P.nsets = (int64_t)pb.aset.size();
  P.nmaps = (int64_t)amap_off.size();
  P.nentries = (int64_t)ent.size();
  P.nmapped = nmapped;
  P.hA_slots = (int64_t)hsrc.size();
  P.load_cost = cost;
  // device arrays
  h->index = std::make_shared<HopIndex>();
  h->index->prod = ProdH;
  auto up = [&](const void *src, size_t bytes, void **dst) -> qs_status {
    QS_TRY(qs_malloc(ctx, dst, std::max<size_t>(bytes, 16)));
    h->index->buf.push_back(*dst);
    if (bytes) QS_TRY(qs_h2d(ctx, *dst, src, bytes));
    return QS_OK;
  };
  void *p_slice, *p_eptr, *p_ent, *p_amap, *p_aset, *p_hA, *p_inv;
  QS_TRY(up(slices.data(), slices.size() * sizeof(int4), &p_slice));
  QS_TRY(up(eptr.data(), eptr.size() * sizeof(uint2), &p_eptr));
  // entry batches are loaded without bounds tests: zero entries past the end
  ent.resize(ent.size() + QS_PROD_ENT_PAD, make_uint4(0u, 0u, 0u, 0u));
  QS_TRY(up(ent.data(), ent.size() * sizeof(uint4), &p_ent));
  QS_TRY(up(amap.data(), amap.size() * 4, &p_amap));
  QS_TRY(up(asets.data(), asets.size() * sizeof(int4), &p_aset));
  QS_TRY(up(hsrc.data(), hsrc.size() * 4, &p_hA));
  QS_TRY(up(inv.data(), (size_t)n * 4, &p_inv));
  SellD &A = h->view;
  A = SellD();
  A.kind = 1;
  A.nrows = n;
  A.nslices = (int64_t)slices.size();
  A.nw = nw;
  A.pslice = (const int4 *)p_slice;
  A.pent_ptr = (const uint2 *)p_eptr;
  A.pent = (const uint4 *)p_ent;
  A.amap = (const int32_t *)p_amap;
  A.aset = (const int4 *)p_aset;
  A.hA = (const uint32_t *)p_hA;
  h->d_inv = (const int32_t *)p_inv;
  h->nslices = A.nslices;
  h->nnz_offdiag = nnz;
  h->nnz_eff = nnz + dhits;
  h->nnz_stored = (int64_t)(ent.size() + hsrc.size());
  h->nblocks = 0;
  h->block_hits = 0;
  h->ell_slots = (int64_t)hsrc.size();
  h->shift_min = 0;
  return QS_OK;
}

static void prod_stream_bytes(qs_hop *h) {
  const ProdHost &P = *h->index->prod;
  const int64_t n = h->rows->n;
  // slice headers, entry offsets and entries, A-matrices, diagonal; the maps
  // and A-matrices are read from cache by every group that uses them
  h->stream_bytes = (int64_t)(h->nslices * 16 + (P.ngroups + 1) * 8 + P.nentries * 16 +
                              P.hA_slots * 4 + P.nsets * 16 + n * (h->view.real_diag ? 8 : 16));
}

// ---------------------------------------------------------------------------
// entry points
// ---------------------------------------------------------------------------
extern "C" qs_status qs_hop_create_product(qs_ctx *ctx, const qs_op *op, const qs_basis *basis,
                                           const uint8_t *amask, qs_hop **out) {
  if (!ctx || !op || !basis || !amask || !out) return QS_EINVAL;
  *out = nullptr;
  if (op->nf != basis->nf)
    return qs_fail(ctx, QS_EINVAL, "operator has %d modes, the basis %d", op->nf, basis->nf);
  QS_HIP(ctx, hipSetDevice(ctx->device));
  qs_hop *h = new qs_hop();
  h->op = op;
  h->rows = basis;
  h->cols = basis;
  h->same = true;
  qs_status st = prod_build(ctx, h, amask);
  if (st == QS_OK) st = prod_bind_values(ctx, h);
  if (st != QS_OK) {
    qs_hop_destroy(h);
    return st;
  }
  prod_stream_bytes(h);
  h->assembled = 1;
  *out = h;
  return QS_OK;
}

qs_status qs_prod_create_like(qs_ctx *ctx, const qs_op *op, const qs_hop *like, qs_hop **out) {
  qs_hop *h = new qs_hop();
  h->op = op;
  h->rows = like->rows;
  h->cols = like->cols;
  h->same = true;
  h->assembled = 1;
  h->index = like->index;
  h->view = like->view;
  h->view.diag = nullptr;
  h->d_inv = like->d_inv;
  h->nslices = like->nslices;
  h->nnz_eff = like->nnz_eff;
  h->nnz_offdiag = like->nnz_offdiag;
  h->nnz_stored = like->nnz_stored;
  h->ell_slots = like->ell_slots;
  qs_status st = prod_bind_values(ctx, h);
  if (st != QS_OK) {
    qs_hop_destroy(h);
    return st;
  }
  prod_stream_bytes(h);
  *out = h;
  return QS_OK;
}

qs_status qs_prod_to_rows(qs_ctx *ctx, const qs_hop *h, int k, const double2 *x, int64_t ldvx, double2 *xp,
                          int64_t ldvp) {
  const int64_t n = h->rows->n;
  if (n == 0 || k == 0) return QS_OK;
  hipLaunchKernelGGL(k_prod_to_rows, dim3(grid_for(n)), dim3(256), 0, ctx->stream, n, k, h->d_inv, x, ldvx, xp, ldvp);
  QS_HIP(ctx, hipGetLastError());
  return QS_OK;
}

qs_status qs_prod_from_rows(qs_ctx *ctx, const qs_hop *h, int k, const double2 *yp, int64_t ldvp, double2 *y,
                            int64_t ldvy) {
  const int64_t n = h->rows->n;
  if (n == 0 || k == 0) return QS_OK;
  hipLaunchKernelGGL(k_prod_from_rows, dim3(grid_for(n)), dim3(256), 0, ctx->stream, n, k, h->d_inv, yp, ldvp, y,
                     ldvy);
  QS_HIP(ctx, hipGetLastError());
  return QS_OK;
}

template <int KC>
static void launch_prod(qs_ctx *ctx, const qs_hop *h, const double2 *x, int64_t ldvx, double2 *y, int64_t ldvy) {
  const SellD &A = h->view;
  const int64_t want = (A.nslices + 3) / 4;
  if (term_value_bytes(A.nterms) <= (size_t)ctx->term_lds) {
    const size_t shm = prod_lds_bytes(A.nterms, true, KC);
    const void *kern = (const void *)k_prod<KC, true>;
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(want, qs_resident_blocks(ctx, kern, shm)));
    hipLaunchKernelGGL((k_prod<KC, true>), dim3(blocks), dim3(QS_BLOCK), shm, ctx->stream, A, x, 1, ldvx, y, 1, ldvy);
  } else {
    const size_t shm = prod_lds_bytes(A.nterms, false, KC);
    const void *kern = (const void *)k_prod<KC, false>;
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(want, qs_resident_blocks(ctx, kern, shm)));
    hipLaunchKernelGGL((k_prod<KC, false>), dim3(blocks), dim3(QS_BLOCK), shm, ctx->stream, A, x, 1, ldvx, y, 1,
                       ldvy);
  }
}

qs_status qs_prod_apply_rows(qs_ctx *ctx, const qs_hop *h, int k, const double2 *x, int64_t ldvx, double2 *y,
                             int64_t ldvy) {
  if (h->rows->n == 0 || k == 0) return QS_OK;
  for (int v0 = 0; v0 < k;) {
    const int kc = k - v0 >= 4 ? 4 : (k - v0 >= 2 ? 2 : 1);
    if (kc == 4) launch_prod<4>(ctx, h, x + v0 * ldvx, ldvx, y + v0 * ldvy, ldvy);
    else if (kc == 2) launch_prod<2>(ctx, h, x + v0 * ldvx, ldvx, y + v0 * ldvy, ldvy);
    else launch_prod<1>(ctx, h, x + v0 * ldvx, ldvx, y + v0 * ldvy, ldvy);
    QS_HIP(ctx, hipGetLastError());
    v0 += kc;
  }
  return QS_OK;
}

extern "C" qs_status qs_hop_row_order(qs_ctx *ctx, const qs_hop *h, int64_t *canon) {
  if (!ctx || !h || (h->rows->n && !canon)) return QS_EINVAL;
  const int64_t n = h->rows->n;
  if (h->view.kind != 1) {
    for (int64_t r = 0; r < n; ++r) canon[r] = r;
    return QS_OK;
  }
  std::vector<int32_t> v((size_t)n);
  if (n) QS_TRY(qs_d2h(ctx, v.data(), h->d_inv, (size_t)n * 4));
  for (int64_t r = 0; r < n; ++r) canon[r] = v[(size_t)r];
  return QS_OK;
}

extern "C" qs_status qs_hop_apply_rows(qs_ctx *ctx, const qs_hop *h, const double *x, double *y) {
  if (!ctx || !h || (h->rows->n && (!x || !y))) return QS_EINVAL;
  QS_HIP(ctx, hipSetDevice(ctx->device));
  if (h->view.kind != 1) return qs_hop_apply(ctx, h, x, y);
  QS_TRY(qs_prod_apply_rows(ctx, h, 1, (const double2 *)x, h->rows->n, (double2 *)y, h->rows->n));
  if (!ctx->async) QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QS_OK;
}

extern "C" qs_status qs_hop_product_info(const qs_hop *h, int64_t *info, double *load_cost) {
  if (!h || !info) return QS_EINVAL;
  if (h->view.kind != 1 || !h->index || !h->index->prod) return QS_EINVAL;
  const ProdHost &P = *h->index->prod;
  const int64_t v[8] = {P.ngroups, h->nslices, P.nsets, P.nentries, P.nmapped, 0, P.hA_slots, P.nmaps};
  for (int i = 0; i < 8; ++i) info[i] = v[i];
  if (load_cost) *load_cost = P.load_cost;
  return QS_OK;
}
