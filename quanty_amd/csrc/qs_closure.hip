// Closure of a determinant set under an operator on the GPU: the reachable
// space the reference grows by repeated H*psi (Core/BlockLanczos.cpp:5849-5858,
// LuaInterface/ssp.cpp:2265-2316), where every (determinant, term) hit adds its
// target to the set whatever the amplitude.
//
// One round: every (frontier determinant, term) pair whose term acts on the
// determinant (push direction: d & C == V ^ F, target d ^ F) and whose target
// passes the restrictions appends the target to a buffer (atomic counter; the
// order does not matter). The buffer is sorted (radix sort; two stable 64-bit
// passes for two-word determinants), made unique, and targets already in the
// set are dropped. What is left is the next frontier; it is merged into the
// set by sorting the union. The closure is reached when a round adds nothing.
// The image of a set (one round from an empty set) is the output basis of the
// reference's push-form OperatorPsi (Core/OperatorWaveFunction.cpp:1750-2258).
#include "qs_internal.h"
#include "qs_kernels.h"

#include <hipcub/hipcub.hpp>

#include <vector>

namespace {

constexpr int CL_MAXRES = 16;
constexpr int64_t CL_ITEMS = (int64_t)1 << 24;  // (determinant, term) pairs per generation launch

struct ResD {
  int n;
  uint64_t m[CL_MAXRES][2];
  int lo[CL_MAXRES], hi[CL_MAXRES];
};

template <int NW>
struct Key {
  uint64_t w[NW];
};

template <int NW>
__device__ __forceinline__ bool key_less(const uint64_t *a, const uint64_t *b) {
  if (NW == 1) return a[0] < b[0];
  return a[0] < b[0] || (a[0] == b[0] && a[1] < b[1]);
}

template <int NW>
__global__ void k_closure_gen(const uint64_t *__restrict__ F, int64_t f0, int64_t fcnt,
                              const TermD<NW> *__restrict__ terms, int nterms, ResD R, uint64_t *__restrict__ out,
                              unsigned long long *__restrict__ cnt) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= fcnt * nterms) return;
  const int64_t f = f0 + idx / nterms;
  const TermD<NW> &t = terms[idx % nterms];
  uint64_t d[NW];
  bool hit = true;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    d[w] = F[f * NW + w];
    hit = hit && ((d[w] & t.C[w]) == (t.V[w] ^ t.F[w]));
  }
  if (!hit) return;
#pragma unroll
  for (int w = 0; w < NW; ++w) d[w] ^= t.F[w];
  for (int r = 0; r < R.n; ++r) {
    int c = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) c += __popcll(d[w] & R.m[r][w]);
    if (c < R.lo[r] || c > R.hi[r]) return;
  }
  const unsigned long long pos = atomicAdd(cnt, 1ULL);
#pragma unroll
  for (int w = 0; w < NW; ++w) out[pos * NW + w] = d[w];
}

// flag[i] = 1 when sorted key i differs from key i-1 and is absent from the
// sorted set S (binary search); 0 otherwise.
template <int NW>
__global__ void k_new_flags(const uint64_t *__restrict__ a, int64_t m, const uint64_t *__restrict__ S, int64_t ns,
                            unsigned char *__restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint64_t *k = a + i * NW;
  bool fresh = i == 0 || key_less<NW>(a + (i - 1) * NW, k);
  if (fresh && S) {
    int64_t lo = 0, hi = ns;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (key_less<NW>(S + mid * NW, k))
        lo = mid + 1;
      else
        hi = mid;
    }
    if (lo < ns) {
      bool eq = true;
#pragma unroll
      for (int w = 0; w < NW; ++w) eq = eq && S[lo * NW + w] == k[w];
      if (eq) fresh = false;
    }
  }
  flag[i] = fresh ? 1 : 0;
}

__global__ void k_iota(int64_t n, int64_t *__restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

// dst[i] = src word w of entry idx[i]
__global__ void k_gather_word(int64_t n, const uint64_t *__restrict__ src, int nw, int w,
                              const int64_t *__restrict__ idx, uint64_t *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i] * nw + w];
}

__global__ void k_gather_key2(int64_t n, const uint64_t *__restrict__ src, const int64_t *__restrict__ idx,
                              uint64_t *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    dst[2 * i] = src[2 * idx[i]];
    dst[2 * i + 1] = src[2 * idx[i] + 1];
  }
}

inline unsigned grid_of(int64_t n) { return (unsigned)((n + 255) / 256); }

// Device buffers of the closure, released together.
struct Bufs {
  std::vector<void *> list;
  ~Bufs() {
    for (void *p : list) qs_free(p);
  }
  qs_status get(qs_ctx *ctx, size_t bytes, void **p) {
    QS_TRY(qs_malloc(ctx, p, bytes ? bytes : 8));
    list.push_back(*p);
    return QS_OK;
  }
};

// Sorts m keys of NW words from `in` into `out` (ascending); `in` is clobbered
// for NW == 2. Temporary storage from bufs.
template <int NW>
qs_status sort_keys(qs_ctx *ctx, Bufs &bufs, uint64_t *in, uint64_t *out, int64_t m) {
  if (m == 0) return QS_OK;
  if (NW == 1) {
    size_t tb = 0;
    QS_HIP(ctx, hipcub::DeviceRadixSort::SortKeys(nullptr, tb, in, out, m, 0, 64, ctx->stream));
    void *tmp;
    QS_TRY(bufs.get(ctx, tb, &tmp));
    QS_HIP(ctx, hipcub::DeviceRadixSort::SortKeys(tmp, tb, in, out, m, 0, 64, ctx->stream));
    return QS_OK;
  }
  // (hi, lo) order by two stable passes: by the low word, then by the high word
  int64_t *i0, *i1;
  uint64_t *k0, *k1;
  QS_TRY(bufs.get(ctx, (size_t)m * 8, (void **)&i0));
  QS_TRY(bufs.get(ctx, (size_t)m * 8, (void **)&i1));
  QS_TRY(bufs.get(ctx, (size_t)m * 8, (void **)&k0));
  QS_TRY(bufs.get(ctx, (size_t)m * 8, (void **)&k1));
  hipLaunchKernelGGL(k_iota, dim3(grid_of(m)), dim3(256), 0, ctx->stream, m, i0);
  hipLaunchKernelGGL(k_gather_word, dim3(grid_of(m)), dim3(256), 0, ctx->stream, m, (const uint64_t *)in, 2, 1,
                     (const int64_t *)i0, k0);
  size_t tb = 0;
  QS_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, i0, i1, m, 0, 64, ctx->stream));
  void *tmp;
  QS_TRY(bufs.get(ctx, tb, &tmp));
  QS_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, i0, i1, m, 0, 64, ctx->stream));
  hipLaunchKernelGGL(k_gather_word, dim3(grid_of(m)), dim3(256), 0, ctx->stream, m, (const uint64_t *)in, 2, 0,
                     (const int64_t *)i1, k0);
  QS_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, i1, i0, m, 0, 64, ctx->stream));
  hipLaunchKernelGGL(k_gather_key2, dim3(grid_of(m)), dim3(256), 0, ctx->stream, m, (const uint64_t *)in,
                     (const int64_t *)i0, out);
  QS_HIP(ctx, hipGetLastError());
  return QS_OK;
}

// out <- the sorted unique keys of in[0..m) that are not in S; returns the count.
template <int NW>
qs_status sort_new(qs_ctx *ctx, Bufs &bufs, uint64_t *in, int64_t m, const uint64_t *S, int64_t ns, uint64_t *out,
                   int64_t *count) {
  *count = 0;
  if (m == 0) return QS_OK;
  uint64_t *sorted;
  unsigned char *flag;
  int64_t *dn;
  QS_TRY(bufs.get(ctx, (size_t)m * 8 * NW, (void **)&sorted));
  QS_TRY(bufs.get(ctx, (size_t)m, (void **)&flag));
  QS_TRY(bufs.get(ctx, 8, (void **)&dn));
  QS_TRY(sort_keys<NW>(ctx, bufs, in, sorted, m));
  hipLaunchKernelGGL(k_new_flags<NW>, dim3(grid_of(m)), dim3(256), 0, ctx->stream, (const uint64_t *)sorted, m, S, ns,
                     flag);
  QS_HIP(ctx, hipGetLastError());
  const Key<NW> *kin = (const Key<NW> *)sorted;
  Key<NW> *kout = (Key<NW> *)out;
  size_t tb = 0;
  QS_HIP(ctx, hipcub::DeviceSelect::Flagged(nullptr, tb, kin, flag, kout, dn, m, ctx->stream));
  void *tmp;
  QS_TRY(bufs.get(ctx, tb, &tmp));
  QS_HIP(ctx, hipcub::DeviceSelect::Flagged(tmp, tb, kin, flag, kout, dn, m, ctx->stream));
  QS_TRY(qs_d2h(ctx, count, dn, 8));
  QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QS_OK;
}

// image = false: the closure of seed. image = true: one round only, from an
// empty set, so the result is the set of targets of the seed's determinants
// (the seed itself only where a term maps a determinant onto one of them).
template <int NW>
qs_status closure(qs_ctx *ctx, const qs_op *op, const qs_basis *seed, const ResD &R, int64_t max_dim,
                  bool image, qs_basis **out) {
  Bufs bufs;
  const int64_t nt = op->nterms;
  // S: the set so far (sorted); F: the last round's new determinants (sorted)
  int64_t ns = image ? 0 : seed->n;
  uint64_t *S;
  QS_TRY(qs_malloc(ctx, (void **)&S, (size_t)(ns ? ns : 1) * 8 * NW));
  std::vector<void *> owned{S};
  auto release = [&]() {
    for (void *p : owned) qs_free(p);
    owned.clear();
  };
  hipError_t e = ns ? hipMemcpyAsync(S, seed->d_words, (size_t)ns * 8 * NW, hipMemcpyDeviceToDevice, ctx->stream)
                    : hipSuccess;
  if (e != hipSuccess) {
    release();
    return qs_hip_fail(ctx, e, "qs_basis_closure", __FILE__, __LINE__);
  }
  const uint64_t *F = image ? seed->d_words : S;
  int64_t nf = seed->n;
  unsigned long long *dcnt;
  qs_status st = bufs.get(ctx, 8, (void **)&dcnt);
  while (st == QS_OK && nf > 0 && nt > 0) {
    // new determinants of this round, accumulated over chunks of the frontier
    uint64_t *acc = nullptr;
    int64_t nacc = 0, cap_acc = 0;
    const int64_t chunk = std::max<int64_t>(1, CL_ITEMS / nt);
    for (int64_t f0 = 0; st == QS_OK && f0 < nf; f0 += chunk) {
      const int64_t fc = std::min(chunk, nf - f0);
      uint64_t *gen;
      Bufs round;
      st = round.get(ctx, (size_t)(fc * nt) * 8 * NW, (void **)&gen);
      if (st != QS_OK) break;
      e = hipMemsetAsync(dcnt, 0, 8, ctx->stream);
      if (e == hipSuccess) {
        hipLaunchKernelGGL(k_closure_gen<NW>, dim3(grid_of(fc * nt)), dim3(256), 0, ctx->stream, F, f0, fc,
                           (const TermD<NW> *)op->d_terms, (int)nt, R, gen, dcnt);
        e = hipGetLastError();
      }
      unsigned long long m = 0;
      if (e == hipSuccess) e = hipMemcpyAsync(&m, dcnt, 8, hipMemcpyDeviceToHost, ctx->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) {
        st = qs_hip_fail(ctx, e, "qs_basis_closure", __FILE__, __LINE__);
        break;
      }
      // append this chunk's new targets (unique, not in S) to acc, then make acc unique
      uint64_t *fresh;
      int64_t nfresh = 0;
      st = round.get(ctx, (size_t)(m ? m : 1) * 8 * NW, (void **)&fresh);
      if (st == QS_OK) st = sort_new<NW>(ctx, round, gen, (int64_t)m, S, ns, fresh, &nfresh);
      if (st != QS_OK || nfresh == 0) continue;
      uint64_t *cat;
      st = round.get(ctx, (size_t)(nacc + nfresh) * 8 * NW, (void **)&cat);
      if (st != QS_OK) break;
      e = nacc ? hipMemcpyAsync(cat, acc, (size_t)nacc * 8 * NW, hipMemcpyDeviceToDevice, ctx->stream) : hipSuccess;
      if (e == hipSuccess)
        e = hipMemcpyAsync(cat + nacc * NW, fresh, (size_t)nfresh * 8 * NW, hipMemcpyDeviceToDevice, ctx->stream);
      if (e != hipSuccess) {
        st = qs_hip_fail(ctx, e, "qs_basis_closure", __FILE__, __LINE__);
        break;
      }
      if (nacc + nfresh > cap_acc) {
        uint64_t *grown;
        cap_acc = 2 * (nacc + nfresh);
        st = bufs.get(ctx, (size_t)cap_acc * 8 * NW, (void **)&grown);
        if (st != QS_OK) break;
        acc = grown;
      }
      int64_t nu = 0;
      st = sort_new<NW>(ctx, round, cat, nacc + nfresh, nullptr, 0, acc, &nu);
      nacc = nu;
      if (st == QS_OK && ns + nacc > max_dim) st = QS_ENOTCONVERGED;
    }
    if (st != QS_OK) break;
    if (image) {  // the targets of one round are the result
      F = acc;
      ns = nacc;
      break;
    }
    if (nacc == 0) break;
    // S <- sort(S u acc); the next frontier is acc
    uint64_t *cat, *S2;
    st = bufs.get(ctx, (size_t)(ns + nacc) * 8 * NW, (void **)&cat);
    if (st == QS_OK) st = qs_malloc(ctx, (void **)&S2, (size_t)(ns + nacc) * 8 * NW);
    if (st != QS_OK) break;
    owned.push_back(S2);
    e = hipMemcpyAsync(cat, S, (size_t)ns * 8 * NW, hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(cat + ns * NW, acc, (size_t)nacc * 8 * NW, hipMemcpyDeviceToDevice, ctx->stream);
    if (e != hipSuccess) {
      st = qs_hip_fail(ctx, e, "qs_basis_closure", __FILE__, __LINE__);
      break;
    }
    st = sort_keys<NW>(ctx, bufs, cat, S2, ns + nacc);
    if (st != QS_OK) break;
    S = S2;
    ns += nacc;
    F = acc;
    nf = nacc;
  }
  if (st == QS_OK) e = hipStreamSynchronize(ctx->stream);
  else (void)hipStreamSynchronize(ctx->stream);
  if (st == QS_OK && e != hipSuccess) st = qs_hip_fail(ctx, e, "qs_basis_closure", __FILE__, __LINE__);
  if (st == QS_ENOTCONVERGED) {
    release();
    return qs_fail(ctx, QS_ENOTCONVERGED, "qs_basis_closure: the set passed max_dim = %lld", (long long)max_dim);
  }
  if (st != QS_OK) {
    release();
    return st;
  }
  qs_basis *b = nullptr;
  st = qs_basis_from_device_words(ctx, seed->nf, image ? F : S, ns, &b);
  release();
  if (st != QS_OK) return st;
  *out = b;
  return QS_OK;
}

}  // namespace

namespace {
qs_status closure_entry(qs_ctx *ctx, const qs_op *op, const qs_basis *seed, int nres, const uint8_t *masks,
                        const int *lo, const int *hi, int64_t max_dim, bool image, qs_basis **out) {
  if (!ctx || !op || !seed || !out || nres < 0 || (nres > 0 && (!masks || !lo || !hi))) return QS_EINVAL;
  *out = nullptr;
  if (op->nf != seed->nf) return qs_fail(ctx, QS_EINVAL, "operator has %d modes, the seed %d", op->nf, seed->nf);
  if (nres > CL_MAXRES) return qs_fail(ctx, QS_EINVAL, "qs_basis_closure: at most %d restrictions", CL_MAXRES);
  if (!image && seed->n > max_dim)
    return qs_fail(ctx, QS_ENOTCONVERGED, "qs_basis_closure: the seed already has more than max_dim = %lld",
                   (long long)max_dim);
  ResD R{};
  R.n = nres;
  const int nbytes = (seed->nf + 7) / 8;
  for (int r = 0; r < nres; ++r) {
    for (int b = 0; b < nbytes; ++b) R.m[r][b >> 3] |= (uint64_t)masks[r * nbytes + b] << (56 - 8 * (b & 7));
    R.lo[r] = lo[r];
    R.hi[r] = hi[r];
  }
  QS_HIP(ctx, hipSetDevice(ctx->device));
  return seed->nw == 1 ? closure<1>(ctx, op, seed, R, max_dim, image, out)
                       : closure<2>(ctx, op, seed, R, max_dim, image, out);
}
}  // namespace

extern "C" qs_status qs_basis_closure(qs_ctx *ctx, const qs_op *op, const qs_basis *seed, int nres,
                                      const uint8_t *masks, const int *lo, const int *hi, int64_t max_dim,
                                      qs_basis **out) {
  return closure_entry(ctx, op, seed, nres, masks, lo, hi, max_dim, false, out);
}

extern "C" qs_status qs_basis_image(qs_ctx *ctx, const qs_op *op, const qs_basis *seed, int nres,
                                    const uint8_t *masks, const int *lo, const int *hi, int64_t max_dim,
                                    qs_basis **out) {
  return closure_entry(ctx, op, seed, nres, masks, lo, hi, max_dim, true, out);
}
