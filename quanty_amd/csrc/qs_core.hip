// Context, device memory, and restricted-occupation bases (enumeration by
// exact ranking, lookups by rank table or binary search).
#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstring>
#include <iterator>
#include <map>

#include <mutex>
#include <unordered_map>
#include "qs_arena.h"
#include "qs_internal.h"

// ---------------------------------------------------------------------------
// errors and memory
// ---------------------------------------------------------------------------
qs_status qs_fail(qs_ctx *ctx, qs_status st, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  return st;
}

qs_status qs_hip_fail(qs_ctx *ctx, hipError_t e, const char *what, const char *file, int line) {
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return qs_fail(ctx, QS_ENOMEM_DEVICE, "%s: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  return qs_fail(ctx, QS_EHIP, "%s: %s (%s:%d)", what, hipGetErrorString(e), file, line);
}

// Device memory. The library's blocks are carved from chunks of device
// memory that stay with the process (QsArena, qs_arena.h): the driver clears
// memory when it hands it out, tens of milliseconds per GB, and other threads'
// allocations wait meanwhile, so a serving stream that returned and re-took its
// multi-GB Krylov bases and assembled operators per request would spend
// seconds in the driver. A chunk holds at least QS_CHUNK bytes (a larger
// request gets a chunk of its own size) and is only handed back to the driver
// when it is wholly free and a new chunk would otherwise leave less than
// QS_FREE_RESERVE of the device to other allocators, or by qs_trim_memory.
// Allocating and freeing one large block once reserves that much memory for
// later blocks of any size.
#define QS_FREE_RESERVE ((size_t)8 << 30)
#define QS_CHUNK ((size_t)4 << 30)

namespace {
std::mutex g_alloc_mu;                      // guards the maps only; driver calls run outside it
std::map<int, QsArena> g_arena;             // device -> its chunks and blocks
std::unordered_map<void *, int> g_owner;    // live block -> device
}  // namespace

// QS_ALLOC_TRACE=1 in the environment: every chunk taken from or returned to
// the driver is reported on stderr (monotonic ns, bytes, duration).
static bool alloc_trace() {
  static const bool on = [] {
    const char *v = getenv("QS_ALLOC_TRACE");
    return v && v[0] == '1';
  }();
  return on;
}

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void driver_free(void *p, size_t bytes) {
  const double t0 = alloc_trace() ? now_s() : 0.0;
  (void)hipFree(p);
  if (alloc_trace()) fprintf(stderr, "[qs alloc] %.0f hipFree %zu B %.3f ms\n", 1e9 * t0, bytes, 1e3 * (now_s() - t0));
}

static hipError_t driver_malloc(void **p, size_t bytes) {
  const double t0 = alloc_trace() ? now_s() : 0.0;
  const hipError_t e = hipMalloc(p, bytes);
  if (alloc_trace()) fprintf(stderr, "[qs alloc] %.0f hipMalloc %zu B %.3f ms\n", 1e9 * t0, bytes, 1e3 * (now_s() - t0));
  return e;
}

// Wholly free chunks of `dev` back to the driver until at most keep bytes stay free.
static void release_chunks(int dev, size_t keep) {
  std::vector<std::pair<uintptr_t, size_t>> out;
  {
    std::lock_guard<std::mutex> lock(g_alloc_mu);
    out = g_arena[dev].release_free_chunks(keep);
  }
  for (auto &c : out) driver_free((void *)c.first, c.second);
}

qs_status qs_malloc(qs_ctx *ctx, void **p, size_t bytes) {
  *p = nullptr;
  const size_t want = QsArena::round(bytes);
  const size_t gran = (size_t)2 << 20;
  for (int attempt = 0; attempt < 8; ++attempt) {
    {
      std::lock_guard<std::mutex> lock(g_alloc_mu);
      QsArena &ar = g_arena[ctx->device];
      const uintptr_t a = ctx->alloc_top ? ar.take_top(want) : ar.take(want);
      if (a) {
        *p = (void *)a;
        g_owner[*p] = ctx->device;
        return QS_OK;
      }
    }
    // a new chunk, keeping QS_FREE_RESERVE of the device for other allocators
    const size_t exact = (want + gran - 1) / gran * gran;
    size_t chunk = std::max(exact, QS_CHUNK);
    size_t dev_free = 0, dev_total = 0;
    if (hipMemGetInfo(&dev_free, &dev_total) == hipSuccess) {
      if (dev_free < chunk + QS_FREE_RESERVE) chunk = exact;
      if (dev_free < chunk + QS_FREE_RESERVE) {
        release_chunks(ctx->device, 0);
        (void)hipMemGetInfo(&dev_free, &dev_total);
      }
    }
    (void)hipGetLastError();
    void *base = nullptr;
    hipError_t e = driver_malloc(&base, chunk);
    if (e != hipSuccess && chunk > exact) {
      (void)hipGetLastError();
      chunk = exact;
      e = driver_malloc(&base, chunk);
    }
    if (e != hipSuccess) {
      (void)hipGetLastError();
      release_chunks(ctx->device, 0);
      e = driver_malloc(&base, chunk);
    }
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return qs_fail(ctx, QS_ENOMEM_DEVICE, "hipMalloc of %zu bytes failed: %s", chunk, hipGetErrorString(e));
    }
    std::lock_guard<std::mutex> lock(g_alloc_mu);
    QsArena &ar = g_arena[ctx->device];
    ar.add_chunk((uintptr_t)base, chunk);
    const uintptr_t a = ctx->alloc_top ? ar.take_top(want) : ar.take(want);
    if (a) {
      *p = (void *)a;
      g_owner[*p] = ctx->device;
      return QS_OK;
    }
    // another thread took the new chunk first: try again
  }
  return qs_fail(ctx, QS_ENOMEM_DEVICE, "could not place a block of %zu bytes", want);
}

void qs_free(void *p) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lock(g_alloc_mu);
    auto it = g_owner.find(p);
    if (it != g_owner.end()) {
      g_arena[it->second].give((uintptr_t)p);
      g_owner.erase(it);
      return;
    }
  }
  (void)hipFree(p);  // not from qs_malloc
}

extern "C" qs_status qs_trim_memory(qs_ctx *ctx, size_t keep_bytes) {
  if (!ctx) return QS_EINVAL;
  QS_HIP(ctx, hipSetDevice(ctx->device));
  release_chunks(ctx->device, keep_bytes);
  return QS_OK;
}

extern "C" qs_status qs_memory_stats(qs_ctx *ctx, size_t *live_bytes, size_t *cached_bytes) {
  if (!ctx) return QS_EINVAL;
  std::lock_guard<std::mutex> lock(g_alloc_mu);
  const QsArena &ar = g_arena[ctx->device];
  if (live_bytes) *live_bytes = ar.live_bytes();
  if (cached_bytes) *cached_bytes = ar.free_bytes();
  return QS_OK;
}

qs_status qs_scratch(qs_ctx *ctx, size_t bytes, void **p) {
  if (bytes > ctx->scratch_bytes) {
    if (ctx->scratch) {
      QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
      qs_free(ctx->scratch);
      ctx->scratch = nullptr;
      ctx->scratch_bytes = 0;
    }
    size_t want = std::max(bytes, (size_t)1 << 20);
    QS_TRY(qs_malloc(ctx, &ctx->scratch, want));
    ctx->scratch_bytes = want;
  }
  *p = ctx->scratch;
  return QS_OK;
}

int64_t qs_resident_blocks(qs_ctx *ctx, const void *kernel, size_t shm) {
  const auto key = std::make_pair(kernel, shm);
  auto it = ctx->resident.find(key);
  int per_cu;
  if (it != ctx->resident.end()) {
    per_cu = it->second;
  } else {
    per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, shm) != hipSuccess || per_cu < 1) {
      (void)hipGetLastError();
      per_cu = 1;
    }
    ctx->resident[key] = per_cu;
  }
  return (int64_t)per_cu * ctx->num_cu;
}

extern "C" qs_status qs_ctx_create(int device, qs_ctx **out) {
  if (!out) return QS_EINVAL;
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) return QS_EHIP;
  if (device < 0 || device >= ndev) return QS_EINVAL;
  qs_ctx *ctx = new qs_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete ctx;
    return QS_EHIP;
  }
  // default-priority stream (no priority other than the default is used)
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return QS_EHIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cu = prop.multiProcessorCount;
  *out = ctx;
  return QS_OK;
}

// pinned staging: at most this many bytes per chunk
#define QS_PINNED_MAX ((size_t)32 << 20)

static qs_status pinned_get(qs_ctx *ctx, size_t bytes) {
  if (ctx->pinned_bytes >= bytes) return QS_OK;
  if (ctx->pinned) {
    (void)hipHostFree(ctx->pinned);
    ctx->pinned = nullptr;
    ctx->pinned_bytes = 0;
  }
  size_t want = (size_t)1 << 20;
  while (want < bytes) want <<= 1;
  QS_HIP(ctx, hipHostMalloc(&ctx->pinned, want, hipHostMallocDefault));
  ctx->pinned_bytes = want;
  return QS_OK;
}

qs_status qs_h2d(qs_ctx *ctx, void *dst, const void *src, size_t bytes) {
  for (size_t off = 0; off < bytes;) {
    const size_t c = std::min(bytes - off, QS_PINNED_MAX);
    QS_TRY(pinned_get(ctx, c));
    QS_HIP(ctx, hipStreamSynchronize(ctx->stream));  // the buffer's previous copy is done
    memcpy(ctx->pinned, (const char *)src + off, c);
    QS_HIP(ctx, hipMemcpyAsync((char *)dst + off, ctx->pinned, c, hipMemcpyHostToDevice, ctx->stream));
    off += c;
  }
  QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QS_OK;
}

qs_status qs_d2h(qs_ctx *ctx, void *dst, const void *src, size_t bytes) {
  for (size_t off = 0; off < bytes;) {
    const size_t c = std::min(bytes - off, QS_PINNED_MAX);
    QS_TRY(pinned_get(ctx, c));
    QS_HIP(ctx, hipMemcpyAsync(ctx->pinned, (const char *)src + off, c, hipMemcpyDeviceToHost, ctx->stream));
    QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    memcpy((char *)dst + off, ctx->pinned, c);
    off += c;
  }
  if (bytes == 0) QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QS_OK;
}

extern "C" qs_status qs_ctx_destroy(qs_ctx *ctx) {
  if (!ctx) return QS_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  if (ctx->scratch) qs_free(ctx->scratch);
  qs_blas_release(ctx);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return QS_OK;
}

extern "C" const char *qs_last_error(const qs_ctx *ctx) { return ctx ? ctx->err.c_str() : "null ctx"; }

extern "C" qs_status qs_dev_alloc(qs_ctx *ctx, size_t bytes, void **ptr) {
  if (!ctx || !ptr) return QS_EINVAL;
  return qs_malloc(ctx, ptr, bytes);
}
extern "C" qs_status qs_dev_free(qs_ctx *ctx, void *ptr) {
  if (!ctx) return QS_EINVAL;
  qs_free(ptr);
  return QS_OK;
}
extern "C" qs_status qs_memcpy_h2d(qs_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return QS_EINVAL;
  return qs_h2d(ctx, dst, src, bytes);
}
extern "C" qs_status qs_memcpy_d2h(qs_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return QS_EINVAL;
  return qs_d2h(ctx, dst, src, bytes);
}
extern "C" void *qs_ctx_stream(qs_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

extern "C" qs_status qs_ctx_set_async(qs_ctx *ctx, int async) {
  if (!ctx) return QS_EINVAL;
  ctx->async = async != 0;
  return QS_OK;
}

extern "C" qs_status qs_ctx_set_param(qs_ctx *ctx, int param, int64_t value) {
  if (!ctx) return QS_EINVAL;
  switch (param) {
    case QS_PARAM_SHIFT_MIN:
      if (value < 1 || value > 65) return qs_fail(ctx, QS_EINVAL, "QS_PARAM_SHIFT_MIN must be in 1..65");
      ctx->shift_min = (int)value;
      return QS_OK;
    case QS_PARAM_ELL_PACK:
      if (value != 0 && value != 1) return qs_fail(ctx, QS_EINVAL, "QS_PARAM_ELL_PACK must be 0 or 1");
      ctx->ell_pack = (int)value;
      return QS_OK;
    case QS_PARAM_BATCH_GRAPH:
      if (value != 0 && value != 1) return qs_fail(ctx, QS_EINVAL, "QS_PARAM_BATCH_GRAPH must be 0 or 1");
      ctx->batch_graph = (int)value;
      return QS_OK;
    case QS_PARAM_TERM_LDS:
      if (value < 0 || value > QS_TERM_LDS_MAX)
        return qs_fail(ctx, QS_EINVAL, "QS_PARAM_TERM_LDS must be in 0..%d", QS_TERM_LDS_MAX);
      ctx->term_lds = (int)value;
      return QS_OK;
    case QS_PARAM_ELL_ORDER:
      if (value != 0 && value != 1) return qs_fail(ctx, QS_EINVAL, "QS_PARAM_ELL_ORDER must be 0 or 1");
      ctx->ell_order = (int)value;
      return QS_OK;
    case QS_PARAM_ALLOC_TOP:
      if (value != 0 && value != 1) return qs_fail(ctx, QS_EINVAL, "QS_PARAM_ALLOC_TOP must be 0 or 1");
      ctx->alloc_top = (int)value;
      return QS_OK;
    case QS_PARAM_HETRD:
      if (value != 0 && value != 1) return qs_fail(ctx, QS_EINVAL, "QS_PARAM_HETRD must be 0 or 1");
      ctx->hetrd = (int)value;
      return QS_OK;
    case QS_PARAM_TRIDIAG_DC:
      if (value < 1) return qs_fail(ctx, QS_EINVAL, "QS_PARAM_TRIDIAG_DC must be >= 1");
      ctx->tridiag_dc = value;
      return QS_OK;
    case QS_PARAM_WAIT_SPINS:
      if (value < 0 || value > ((int64_t)1 << 40)) return qs_fail(ctx, QS_EINVAL, "QS_PARAM_WAIT_SPINS must be in [0, 2^40]");
      ctx->wait_spins = value;
      return QS_OK;
    default:
      return qs_fail(ctx, QS_EINVAL, "unknown parameter %d", param);
  }
}

extern "C" qs_status qs_ctx_get_param(qs_ctx *ctx, int param, int64_t *value) {
  if (!ctx || !value) return QS_EINVAL;
  switch (param) {
    case QS_PARAM_SHIFT_MIN:
      *value = ctx->shift_min;
      return QS_OK;
    case QS_PARAM_ELL_PACK:
      *value = ctx->ell_pack;
      return QS_OK;
    case QS_PARAM_BATCH_GRAPH:
      *value = ctx->batch_graph;
      return QS_OK;
    case QS_PARAM_TERM_LDS:
      *value = ctx->term_lds;
      return QS_OK;
    case QS_PARAM_ELL_ORDER:
      *value = ctx->ell_order;
      return QS_OK;
    case QS_PARAM_ALLOC_TOP:
      *value = ctx->alloc_top;
      return QS_OK;
    case QS_PARAM_HETRD:
      *value = ctx->hetrd;
      return QS_OK;
    case QS_PARAM_TRIDIAG_DC:
      *value = ctx->tridiag_dc;
      return QS_OK;
    case QS_PARAM_WAIT_SPINS:
      *value = ctx->wait_spins;
      return QS_OK;
    default:
      return qs_fail(ctx, QS_EINVAL, "unknown parameter %d", param);
  }
}

extern "C" qs_status qs_sync(qs_ctx *ctx) {
  if (!ctx) return QS_EINVAL;
  QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QS_OK;
}

// ---------------------------------------------------------------------------
// rank tables (host)
// ---------------------------------------------------------------------------
namespace {

// Determinant counts saturate at INT64_MAX: a sector can hold more than 2^63
// determinants (68 modes, 34 particles), and a wrapped count would pass every
// size check as a negative number.
inline int64_t sat_add(int64_t a, int64_t b) {
  int64_t r;
  return __builtin_add_overflow(a, b, &r) ? INT64_MAX : r;
}
inline int64_t sat_mul(int64_t a, int64_t b) {
  int64_t r;
  return __builtin_mul_overflow(a, b, &r) ? INT64_MAX : r;
}

struct StateKey {
  std::vector<int16_t> c;
  bool operator<(const StateKey &o) const { return c < o.c; }
};

// The same table for up to 8 restrictions with bounds <= 255: a state (the
// occupation count inside each mask) is packed into one byte per restriction
// of a uint64, and the byte values of a chunk are grouped by the counts they
// add, so each state meets a handful of distinct successors instead of 256.
// State ids follow the packed keys in ascending order.
bool rank_table_packed(int nf, int nres, const uint8_t *masks, const int *lo, const int *hi, RankTable &rt,
                       std::string &err, bool count_only) {
  const int nbytes = (nf + 7) / 8;
  const int nch = nbytes;
  struct Group {
    uint64_t inc;
    std::vector<int> vs;  // ascending byte values with this increment
  };
  std::vector<std::vector<Group>> groups(nch);
  std::vector<std::vector<int>> gidx(nch, std::vector<int>(256, -1));
  for (int c = 0; c < nch; ++c) {
    const int rem = nf - 8 * c;
    const uint32_t forbid = rem >= 8 ? 0u : ((1u << (8 - rem)) - 1u);
    std::map<uint64_t, int> where;
    for (int v = 0; v < 256; ++v) {
      if (v & forbid) continue;
      uint64_t inc = 0;
      for (int r = 0; r < nres; ++r)
        inc |= (uint64_t)__builtin_popcount((unsigned)(masks[r * nbytes + c] & v)) << (8 * r);
      auto it = where.find(inc);
      int g;
      if (it == where.end()) {
        g = (int)groups[c].size();
        where[inc] = g;
        groups[c].push_back({inc, {}});
      } else {
        g = it->second;
      }
      groups[c][g].vs.push_back(v);
      gidx[c][v] = g;
    }
  }
  // successor of a packed state, or false when a count passes its upper bound
  auto step = [&](uint64_t s, uint64_t inc, uint64_t &t) {
    t = s + inc;  // bytes never carry: every count stays <= 255
    for (int r = 0; r < nres; ++r)
      if ((int)((t >> (8 * r)) & 0xFF) > hi[r]) return false;
    return true;
  };
  std::vector<std::vector<uint64_t>> keys(nch + 1);
  keys[0].push_back(0);
  for (int c = 0; c < nch; ++c) {
    std::vector<uint64_t> nxt;
    for (uint64_t st : keys[c])
      for (const Group &g : groups[c]) {
        uint64_t t;
        if (step(st, g.inc, t)) nxt.push_back(t);
      }
    std::sort(nxt.begin(), nxt.end());
    nxt.erase(std::unique(nxt.begin(), nxt.end()), nxt.end());
    if (nxt.size() > 2000000) {
      err = "restriction state space too large";
      return false;
    }
    keys[c + 1] = std::move(nxt);
  }
  // completions per state, backward
  std::vector<std::vector<int64_t>> cnt(nch + 1);
  auto find = [&](int c, uint64_t t) -> int64_t {
    auto it = std::lower_bound(keys[c].begin(), keys[c].end(), t);
    return (it != keys[c].end() && *it == t) ? (int64_t)(it - keys[c].begin()) : -1;
  };
  cnt[nch].resize(keys[nch].size());
  for (size_t i = 0; i < keys[nch].size(); ++i) {
    bool acc = true;
    for (int r = 0; r < nres; ++r) {
      const int v = (int)((keys[nch][i] >> (8 * r)) & 0xFF);
      if (v < lo[r] || v > hi[r]) acc = false;
    }
    cnt[nch][i] = acc ? 1 : 0;
  }
  for (int c = nch - 1; c >= 0; --c) {
    cnt[c].assign(keys[c].size(), 0);
    for (size_t i = 0; i < keys[c].size(); ++i) {
      int64_t tot = 0;
      for (const Group &g : groups[c]) {
        uint64_t t;
        if (!step(keys[c][i], g.inc, t)) continue;
        const int64_t j = find(c + 1, t);
        if (j >= 0) tot = sat_add(tot, sat_mul(cnt[c + 1][j], (int64_t)g.vs.size()));
      }
      cnt[c][i] = tot;
    }
  }
  rt.total = cnt[0][0];
  if (count_only) return true;
  if (rt.total >= ((int64_t)1 << 32) - 1) {
    err = "basis larger than 2^32-2 determinants";
    return false;
  }
  // ids of viable states (nonzero completions), ascending keys
  std::vector<std::vector<int64_t>> id(nch + 1);
  uint32_t smax = 1;
  for (int c = 0; c <= nch; ++c) {
    id[c].assign(keys[c].size(), -1);
    uint32_t k = 0;
    for (size_t i = 0; i < keys[c].size(); ++i)
      if (cnt[c][i] > 0) id[c][i] = k++;
    if (c < nch) smax = std::max(smax, k);
  }
  rt.nchunks = nch;
  rt.smax = smax;
  rt.next.assign((size_t)nch * smax * 256, QS_DEAD);
  rt.less.assign((size_t)nch * smax * 256, 0);
  for (int c = 0; c < nch; ++c) {
    const size_t ng = groups[c].size();
    std::vector<int64_t> gc(ng);
    std::vector<uint32_t> gn(ng);
    for (size_t i = 0; i < keys[c].size(); ++i) {
      if (id[c][i] < 0) continue;
      for (size_t g = 0; g < ng; ++g) {
        gc[g] = 0;
        gn[g] = QS_DEAD;
        uint64_t t;
        if (!step(keys[c][i], groups[c][g].inc, t)) continue;
        const int64_t j = find(c + 1, t);
        if (j < 0 || cnt[c + 1][j] == 0) continue;
        gc[g] = cnt[c + 1][j];
        gn[g] = c + 1 == nch ? 0u : (uint32_t)id[c + 1][j];  // 0 = accepting at the last chunk
      }
      const size_t base = ((size_t)c * smax + (size_t)id[c][i]) * 256;
      int64_t run = 0;
      for (int v = 0; v < 256; ++v) {
        rt.less[base + v] = (uint32_t)run;
        const int g = gidx[c][v];
        if (g < 0 || gc[g] == 0) continue;
        rt.next[base + v] = gn[g];
        run += gc[g];
      }
    }
  }
  return true;
}

}  // namespace

qs_status qs_build_rank_table(int nf, int nres, const uint8_t *masks, const int *lo,
                              const int *hi, RankTable &rt, std::string &err, bool count_only) {
  if (nres >= 0 && nres <= 8 && nf > 0 && nf <= 128) {
    // no count exceeds 128 (nf <= 128), so bounds are clamped into a byte
    int lo8[8], hi8[8];
    for (int r = 0; r < nres; ++r) {
      lo8[r] = std::max(0, std::min(lo[r], 129));
      hi8[r] = std::min(hi[r], 128);
    }
    if (!rank_table_packed(nf, nres, masks, lo8, hi8, rt, err, count_only)) return QS_EINVAL;
    return QS_OK;
  }
  if (nf <= 0 || nf > 128) {
    err = "nf must be in 1..128";
    return QS_EINVAL;
  }
  if (nres < 0 || nres > 64) {
    err = "at most 64 restrictions";
    return QS_EINVAL;
  }
  const int nbytes = (nf + 7) / 8;
  const int nch = nbytes;
  // per chunk: list of valid byte values and their per-restriction counts
  std::vector<std::vector<int>> vals(nch);
  for (int c = 0; c < nch; ++c) {
    int rem = nf - 8 * c;  // modes present in this chunk
    uint32_t forbid = rem >= 8 ? 0u : ((1u << (8 - rem)) - 1u);
    for (int v = 0; v < 256; ++v)
      if ((v & forbid) == 0) vals[c].push_back(v);
  }
  auto count = [&](int r, int c, int v) {
    return __builtin_popcount((unsigned)(masks[r * nbytes + c] & v));
  };
  // forward reachable states per boundary (counts clamped to hi+1 = overflow => dropped)
  std::vector<std::map<StateKey, int64_t>> bnd(nch + 1);
  StateKey s0;
  s0.c.assign(nres, 0);
  bnd[0][s0] = 0;
  for (int c = 0; c < nch; ++c) {
    for (auto &kv : bnd[c]) {
      for (int v : vals[c]) {
        StateKey t = kv.first;
        bool ok = true;
        for (int r = 0; r < nres && ok; ++r) {
          t.c[r] = (int16_t)(t.c[r] + count(r, c, v));
          if (t.c[r] > hi[r]) ok = false;
        }
        if (ok) bnd[c + 1].emplace(t, 0);
      }
    }
    if (bnd[c + 1].size() > 2000000) {
      err = "restriction state space too large";
      return QS_EINVAL;
    }
  }
  // completions, backward
  for (auto &kv : bnd[nch]) {
    bool acc = true;
    for (int r = 0; r < nres; ++r)
      if (kv.first.c[r] < lo[r] || kv.first.c[r] > hi[r]) acc = false;
    kv.second = acc ? 1 : 0;
  }
  for (int c = nch - 1; c >= 0; --c) {
    for (auto &kv : bnd[c]) {
      int64_t tot = 0;
      for (int v : vals[c]) {
        StateKey t = kv.first;
        bool ok = true;
        for (int r = 0; r < nres && ok; ++r) {
          t.c[r] = (int16_t)(t.c[r] + count(r, c, v));
          if (t.c[r] > hi[r]) ok = false;
        }
        if (!ok) continue;
        auto it = bnd[c + 1].find(t);
        if (it != bnd[c + 1].end()) tot = sat_add(tot, it->second);
      }
      kv.second = tot;
    }
  }
  rt.total = bnd[0].begin()->second;
  if (count_only) return QS_OK;
  if (rt.total >= ((int64_t)1 << 32) - 1) {
    err = "basis larger than 2^32-2 determinants";
    return QS_EINVAL;
  }
  // ids of viable states per boundary
  std::vector<std::map<StateKey, uint32_t>> ids(nch + 1);
  uint32_t smax = 1;
  for (int c = 0; c < nch; ++c) {
    uint32_t k = 0;
    for (auto &kv : bnd[c])
      if (kv.second > 0) ids[c][kv.first] = k++;
    smax = std::max(smax, k);
  }
  rt.nchunks = nch;
  rt.smax = smax;
  rt.next.assign((size_t)nch * smax * 256, QS_DEAD);
  rt.less.assign((size_t)nch * smax * 256, 0);
  for (int c = 0; c < nch; ++c) {
    for (auto &kv : ids[c]) {
      const uint32_t sid = kv.second;
      size_t base = ((size_t)c * smax + sid) * 256;
      int64_t run = 0;
      for (int v = 0; v < 256; ++v) {
        rt.less[base + v] = (uint32_t)run;
        if ((std::find(vals[c].begin(), vals[c].end(), v) == vals[c].end())) continue;
        StateKey t = kv.first;
        bool ok = true;
        for (int r = 0; r < nres && ok; ++r) {
          t.c[r] = (int16_t)(t.c[r] + count(r, c, v));
          if (t.c[r] > hi[r]) ok = false;
        }
        if (!ok) continue;
        auto it = bnd[c + 1].find(t);
        if (it == bnd[c + 1].end() || it->second == 0) continue;
        if (c + 1 == nch) {
          rt.next[base + v] = 0;  // accepting
        } else {
          rt.next[base + v] = ids[c + 1][t];
        }
        run += it->second;
      }
    }
  }
  return QS_OK;
}

// ---------------------------------------------------------------------------
// device kernels: unrank (enumeration) and lookup
// ---------------------------------------------------------------------------
template <int NW>
__global__ void k_unrank(const uint2 *__restrict__ tbl, uint32_t smax, int nch, int64_t n,
                         uint64_t *__restrict__ out) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  uint32_t rem = (uint32_t)k;
  uint32_t s = 0;
  uint64_t w[NW];
#pragma unroll
  for (int i = 0; i < NW; ++i) w[i] = 0;
  for (int c = 0; c < nch; ++c) {
    const uint2 *row = tbl + ((size_t)c * smax + s) * 256;
    int lo = 0, hi = 256;
    while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (row[mid].y <= rem) lo = mid; else hi = mid;
    }
    uint2 e = row[lo];
    rem -= e.y;
    s = e.x;
    w[c >> 3] |= (uint64_t)lo << (56 - 8 * (c & 7));
  }
#pragma unroll
  for (int i = 0; i < NW; ++i) out[k * NW + i] = w[i];
}

template <int NW>
__device__ __forceinline__ int64_t rank_lookup(const uint64_t *w, const uint2 *__restrict__ tbl,
                                               uint32_t smax, int nch) {
  uint32_t s = 0, r = 0;
  for (int c = 0; c < nch; ++c) {
    uint32_t v = (uint32_t)(w[c >> 3] >> (56 - 8 * (c & 7))) & 0xFFu;
    uint2 e = tbl[((size_t)c * smax + s) * 256 + v];
    if (e.x == QS_DEAD) return -1;
    r += e.y;
    s = e.x;
  }
  return r;
}

template <int NW>
__device__ __forceinline__ bool det_less(const uint64_t *a, const uint64_t *b) {
  if (NW == 1) return a[0] < b[0];
  return a[0] < b[0] || (a[0] == b[0] && a[1] < b[1]);
}

template <int NW>
__device__ __forceinline__ int64_t bsearch_lookup(const uint64_t *w, const uint64_t *__restrict__ words,
                                                  int64_t n) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (det_less<NW>(words + mid * NW, w)) lo = mid + 1; else hi = mid;
  }
  if (lo < n) {
    bool eq = true;
#pragma unroll
    for (int i = 0; i < NW; ++i) eq = eq && (words[lo * NW + i] == w[i]);
    if (eq) return lo;
  }
  return -1;
}

template <int NW>
__global__ void k_index(const uint64_t *__restrict__ q, int64_t nq, const uint2 *__restrict__ tbl,
                        uint32_t smax, int nch, const uint64_t *__restrict__ words, int64_t n,
                        int64_t *__restrict__ idx) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nq) return;
  uint64_t w[NW];
#pragma unroll
  for (int i = 0; i < NW; ++i) w[i] = q[k * NW + i];
  idx[k] = tbl ? rank_lookup<NW>(w, tbl, smax, nch) : bsearch_lookup<NW>(w, words, n);
}

// ---------------------------------------------------------------------------
// basis API
// ---------------------------------------------------------------------------
static qs_status basis_alloc_words(qs_ctx *ctx, qs_basis *b) {
  return qs_malloc(ctx, (void **)&b->d_words, (size_t)std::max<int64_t>(b->n, 1) * b->nw * 8);
}

extern "C" qs_status qs_basis_count(qs_ctx *ctx, int nf, int nres, const uint8_t *masks,
                                    const int *lo, const int *hi, int64_t *n) {
  RankTable rt;
  std::string err;
  qs_status st = qs_build_rank_table(nf, nres, masks, lo, hi, rt, err, true);
  if (st != QS_OK) return qs_fail(ctx, st, "%s", err.c_str());
  *n = rt.total;
  return QS_OK;
}

extern "C" qs_status qs_basis_enumerate(qs_ctx *ctx, int nf, int nres, const uint8_t *masks,
                                        const int *lo, const int *hi, qs_basis **out) {
  if (!ctx || !out) return QS_EINVAL;
  *out = nullptr;
  RankTable rt;
  std::string err;
  qs_status st = qs_build_rank_table(nf, nres, masks, lo, hi, rt, err, false);
  if (st != QS_OK) return qs_fail(ctx, st, "qs_basis_enumerate: %s", err.c_str());
  QS_HIP(ctx, hipSetDevice(ctx->device));
  qs_basis *b = new qs_basis();
  b->nf = nf;
  b->nw = nf <= 64 ? 1 : 2;
  b->n = rt.total;
  b->has_rank = true;
  b->nchunks = rt.nchunks;
  b->smax = rt.smax;
  b->nres = nres;
  const int nbytes = (nf + 7) / 8;
  b->masks.assign(masks, masks + (size_t)nres * nbytes);
  b->lo.assign(lo, lo + nres);
  b->hi.assign(hi, hi + nres);
  std::vector<uint2> tbl(rt.next.size());
  for (size_t i = 0; i < tbl.size(); ++i) tbl[i] = make_uint2(rt.next[i], rt.less[i]);
  st = qs_malloc(ctx, (void **)&b->d_tbl, tbl.size() * sizeof(uint2));
  if (st == QS_OK) st = basis_alloc_words(ctx, b);
  if (st != QS_OK) {
    qs_basis_destroy(b);
    return st;
  }
  st = qs_h2d(ctx, b->d_tbl, tbl.data(), tbl.size() * sizeof(uint2));
  if (st != QS_OK) {
    qs_basis_destroy(b);
    return st;
  }
  hipError_t e = hipSuccess;
  if (b->n > 0) {
    int64_t blocks = (b->n + 255) / 256;
    if (b->nw == 1)
      hipLaunchKernelGGL(k_unrank<1>, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, b->d_tbl,
                         b->smax, b->nchunks, b->n, b->d_words);
    else
      hipLaunchKernelGGL(k_unrank<2>, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, b->d_tbl,
                         b->smax, b->nchunks, b->n, b->d_words);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    qs_basis_destroy(b);
    return qs_hip_fail(ctx, e, "qs_basis_enumerate", __FILE__, __LINE__);
  }
  *out = b;
  return QS_OK;
}

extern "C" qs_status qs_basis_from_words(qs_ctx *ctx, int nf, const uint64_t *words, int64_t n,
                                         qs_basis **out) {
  if (!ctx || !out || n < 0 || nf <= 0 || nf > 128) return QS_EINVAL;
  *out = nullptr;
  const int nw = nf <= 64 ? 1 : 2;
  std::vector<std::pair<uint64_t, uint64_t>> v((size_t)n);
  const int shift_unused = 64 * nw - nf;  // low bits that must be zero
  for (int64_t i = 0; i < n; ++i) {
    uint64_t hi = words[i * nw];
    uint64_t lo = nw == 2 ? words[i * nw + 1] : 0;
    uint64_t tail = nw == 1 ? hi : lo;
    int unused = nw == 1 ? shift_unused : shift_unused;  // bits of the last word above nf
    if (unused > 0 && unused < 64 && (tail & ((1ULL << unused) - 1)) != 0)
      return qs_fail(ctx, QS_EINVAL, "determinant %lld has bits set beyond mode %d", (long long)i,
                     nf - 1);
    v[(size_t)i] = {hi, lo};
  }
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  QS_HIP(ctx, hipSetDevice(ctx->device));
  qs_basis *b = new qs_basis();
  b->nf = nf;
  b->nw = nw;
  b->n = (int64_t)v.size();
  std::vector<uint64_t> flat((size_t)b->n * nw);
  for (size_t i = 0; i < v.size(); ++i) {
    flat[i * nw] = v[i].first;
    if (nw == 2) flat[i * nw + 1] = v[i].second;
  }
  qs_status st = basis_alloc_words(ctx, b);
  if (st != QS_OK) {
    delete b;
    return st;
  }
  hipError_t e = hipMemcpyAsync(b->d_words, flat.data(), flat.size() * 8, hipMemcpyHostToDevice,
                                ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    qs_basis_destroy(b);
    return qs_hip_fail(ctx, e, "qs_basis_from_words", __FILE__, __LINE__);
  }
  *out = b;
  return QS_OK;
}

qs_status qs_basis_from_device_words(qs_ctx *ctx, int nf, const uint64_t *d_words, int64_t n, qs_basis **out) {
  qs_basis *b = new qs_basis();
  b->nf = nf;
  b->nw = nf <= 64 ? 1 : 2;
  b->n = n;
  qs_status st = basis_alloc_words(ctx, b);
  if (st != QS_OK) {
    delete b;
    return st;
  }
  hipError_t e = n ? hipMemcpyAsync(b->d_words, d_words, (size_t)n * b->nw * 8, hipMemcpyDeviceToDevice, ctx->stream)
                   : hipSuccess;
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    qs_basis_destroy(b);
    return qs_hip_fail(ctx, e, "qs_basis_from_device_words", __FILE__, __LINE__);
  }
  *out = b;
  return QS_OK;
}

extern "C" int64_t qs_basis_size(const qs_basis *b) { return b ? b->n : -1; }
extern "C" int qs_basis_nf(const qs_basis *b) { return b ? b->nf : -1; }
extern "C" int qs_basis_words(const qs_basis *b) { return b ? b->nw : -1; }
extern "C" const uint64_t *qs_basis_device_words(const qs_basis *b) { return b ? b->d_words : nullptr; }

extern "C" qs_status qs_basis_export(qs_ctx *ctx, const qs_basis *b, uint64_t *words) {
  if (!ctx || !b || !words) return QS_EINVAL;
  if (b->n == 0) return QS_OK;
  QS_HIP(ctx, hipMemcpyAsync(words, b->d_words, (size_t)b->n * b->nw * 8, hipMemcpyDeviceToHost,
                             ctx->stream));
  QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QS_OK;
}

extern "C" qs_status qs_basis_index(qs_ctx *ctx, const qs_basis *b, const uint64_t *words, int64_t n,
                                    int64_t *idx) {
  if (!ctx || !b || (n > 0 && (!words || !idx))) return QS_EINVAL;
  if (n == 0) return QS_OK;
  size_t qbytes = (size_t)n * b->nw * 8;
  void *scr;
  QS_TRY(qs_scratch(ctx, qbytes + (size_t)n * 8, &scr));
  uint64_t *dq = (uint64_t *)scr;
  int64_t *di = (int64_t *)((char *)scr + qbytes);
  QS_HIP(ctx, hipMemcpyAsync(dq, words, qbytes, hipMemcpyHostToDevice, ctx->stream));
  int64_t blocks = (n + 255) / 256;
  const uint2 *tbl = b->has_rank ? b->d_tbl : nullptr;
  if (b->nw == 1)
    hipLaunchKernelGGL(k_index<1>, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, dq, n, tbl,
                       b->smax, b->nchunks, b->d_words, b->n, di);
  else
    hipLaunchKernelGGL(k_index<2>, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, dq, n, tbl,
                       b->smax, b->nchunks, b->d_words, b->n, di);
  QS_HIP(ctx, hipGetLastError());
  QS_HIP(ctx, hipMemcpyAsync(idx, di, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream));
  QS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QS_OK;
}

// Index in b of n determinants already in device memory (d_words), into
// d_idx (device); -1 for determinants outside b. Enqueued on ctx->stream.
qs_status qs_basis_index_dev(qs_ctx *ctx, const qs_basis *b, const uint64_t *d_words, int64_t n, int64_t *d_idx) {
  if (n == 0) return QS_OK;
  const int64_t blocks = (n + 255) / 256;
  const uint2 *tbl = b->has_rank ? b->d_tbl : nullptr;
  if (b->nw == 1)
    hipLaunchKernelGGL(k_index<1>, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, d_words, n, tbl, b->smax,
                       b->nchunks, b->d_words, b->n, d_idx);
  else
    hipLaunchKernelGGL(k_index<2>, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, d_words, n, tbl, b->smax,
                       b->nchunks, b->d_words, b->n, d_idx);
  QS_HIP(ctx, hipGetLastError());
  return QS_OK;
}

extern "C" qs_status qs_basis_destroy(qs_basis *b) {
  if (!b) return QS_OK;
  qs_free(b->d_words);
  qs_free(b->d_tbl);
  delete b;
  return QS_OK;
}
