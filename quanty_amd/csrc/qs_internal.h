// Internal declarations shared by the libqs translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <utility>
#include <complex>
#include <memory>
#include <vector>

#include "../../include/qs.h"

#define QS_MAX_WORDS 2
#define QS_MAX_CHUNKS 16          // 128 modes / 8 modes per chunk
#define QS_DEAD 0xFFFFFFFFu        // rank-table successor of an invalid prefix

// Term tables (V_t and P_t) up to this many bytes are staged in LDS by the
// assembled kernels unless the context sets another limit.
#define QS_TERM_LDS_DEFAULT (24 * 1024)
#define QS_TERM_LDS_MAX (64 * 1024)

struct qs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // reusable device scratch (grown on demand, freed with the ctx)
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  int num_cu = 256;
  int async = 0;                   // skip the end-of-call stream sync for device-only results
  // resident workgroups per CU, by (kernel, dynamic LDS bytes)
  std::map<std::pair<const void *, size_t>, int> resident;
  int shift_min = 56;              // QS_PARAM_SHIFT_MIN for operators assembled from now on
  int ell_pack = 1;                // QS_PARAM_ELL_PACK
  int batch_graph = 1;             // QS_PARAM_BATCH_GRAPH
  int term_lds = QS_TERM_LDS_DEFAULT;  // QS_PARAM_TERM_LDS
  int ell_order = 1;               // QS_PARAM_ELL_ORDER
  int alloc_top = 0;               // QS_PARAM_ALLOC_TOP: device blocks from the top of free memory
  int hetrd = 1;                   // QS_PARAM_HETRD: the dense method's own cooperative reduction
  int64_t tridiag_dc = 256;        // QS_PARAM_TRIDIAG_DC: order from which dstedc solves the tridiagonal problem
  int64_t wait_spins = 1 << 22;    // QS_PARAM_WAIT_SPINS: polls a wait of the dense reduction makes before failing
  void *blas = nullptr;            // rocblas handle for the dense eigensolver, made on first use
  void *pinned = nullptr;          // the ctx's own pinned staging buffer (qs_h2d / qs_d2h)
  size_t pinned_bytes = 0;
};

// A basis holding n sorted, unique determinants copied from device memory.
qs_status qs_basis_from_device_words(qs_ctx *ctx, int nf, const uint64_t *d_words, int64_t n, qs_basis **out);

// Eigenvalues of the real symmetric tridiagonal matrix (d[0..n-1], e[0..n-2])
// in ascending order (w) and the projections of nv vectors on its
// eigenvectors: Y holds row i = component i of every vector (n x nv complex,
// row-major); P row k = z_k^T Y for the k-th eigenvalue (qs_dense.cpp).
bool qs_tridiag_project(int n, const double *d, const double *e, int nv, const std::complex<double> *Y, double *w,
                        std::complex<double> *P);

// Eigenvalues (ascending, w) of the real symmetric tridiagonal matrix (d, e)
// and the eigenvectors of its k lowest, Z column j = z_j (n x k, column-major)
// (qs_dense.cpp).
bool qs_tridiag_lowest_vectors(int n, const double *d, const double *e, int k, double *w, double *Z);

// Releases ctx->blas (qs_spectra.hip).
void qs_blas_release(qs_ctx *ctx);

// Exact rank of a determinant among all determinants satisfying the
// restrictions, in ascending numeric order. The modes are cut into chunks of 8
// (one reference byte each). After chunks 0..c-1 the "state" is the vector of
// occupation counts inside every restriction mask; only states that can still
// be completed into a valid determinant get an id. For chunk c, state s and
// byte value v, entry (c, s, v) holds the id of the successor state (or
// QS_DEAD) and the number of valid determinants that share the prefix and have
// a smaller byte at chunk c. rank = sum of the entries along the path.
struct RankTable {
  int nchunks = 0;
  uint32_t smax = 0;               // state stride (max viable states at any boundary)
  std::vector<uint32_t> next;      // nchunks * smax * 256
  std::vector<uint32_t> less;      // nchunks * smax * 256
  int64_t total = 0;               // number of valid determinants
};

struct qs_basis {
  int nf = 0;
  int nw = 1;
  int64_t n = 0;
  uint64_t *d_words = nullptr;     // n * nw, ascending
  bool has_rank = false;
  int nchunks = 0;
  uint32_t smax = 0;
  uint2 *d_tbl = nullptr;          // {next, less} per (chunk, state, byte)
  // restriction description (kept for filtering)
  int nres = 0;
  std::vector<uint8_t> masks;
  std::vector<int> lo, hi;
};

// A ladder string compiled for the pull form. For a row determinant d the
// term contributes iff (d & C) == V; its source determinant is s = d ^ F and
// the fermion sign is (-1)^(popcount(d & P) + sgn).
struct TermHost {
  uint64_t C[QS_MAX_WORDS] = {0, 0};
  uint64_t V[QS_MAX_WORDS] = {0, 0};
  uint64_t F[QS_MAX_WORDS] = {0, 0};
  uint64_t P[QS_MAX_WORDS] = {0, 0};
  uint32_t sgn = 0;
  double re = 0, im = 0;
  int nchange = 0;                 // signed change of particle number
};

template <int NW>
struct TermD {
  uint64_t C[NW], V[NW], F[NW], P[NW];
  double re, im;
  uint32_t sgn;
  uint32_t diag;                   // F == 0
};

struct qs_op {
  int nf = 0;
  int nw = 1;
  int nterms = 0;                  // compiled (non-vanishing) terms
  std::vector<TermHost> h_terms;
  void *d_terms = nullptr;         // TermD<nw>[nterms]
  double *d_val = nullptr;         // V_t per term (complex, fermion-sign constant folded in)
  uint64_t *d_par = nullptr;       // P_t per term (nw words): entry sign = (-1)^popcount(d & P_t)
};

// Device view of an assembled operator. Rows are cut into slices of 64 (one
// wavefront, one lane per row). A slice's off-diagonal entries are split into
//  - shift blocks: one term whose hit lanes l all read column base + l; a
//    16-byte header {base (int32), term, lane mask lo, lane mask hi};
//  - an ELL remainder, column-major inside the slice (entry e of lane l at
//    slice_ptr + 64 e + l): col | term << colbits in one uint32 when both fit,
//    else the column in col[] and the term in tid[].
// The entry sign is (-1)^popcount(d_row & P_term); the diagonal terms are
// pre-summed per row when rows == cols.
struct SellD {
  int64_t nrows = 0;
  int64_t nslices = 0;
  int nw = 1;
  int packed = 0;
  uint32_t colbits = 0;
  int nterms = 0;
  int real_diag = 0;                   // diag holds doubles (every diagonal term is real)
  const uint64_t *rows = nullptr;      // row determinants (nw words each)
  const uint64_t *blk_ptr = nullptr;   // nslices + 1 block offsets
  const uint4 *blk = nullptr;          // shift-block headers
  const uint64_t *slice_ptr = nullptr; // nslices + 1 ELL entry offsets
  const uint32_t *slice_len = nullptr; // ELL positions per slice
  const uint16_t *rowcnt = nullptr;    // ELL entries per row
  const uint32_t *col = nullptr;       // ELL column (or packed column and term)
  const uint16_t *tid = nullptr;       // ELL term when not packed
  const double2 *val = nullptr;        // V_t
  const uint64_t *par = nullptr;       // P_t
  const void *diag = nullptr;          // pre-summed diagonal (rows == cols) or null: double2, or double if real_diag
  // Product layout (kind 1; qs_product.hip). The modes are split into two
  // groups, A and B. Rows are ordered by (B part, A part): a "group" holds the
  // rows of one B configuration b, one row per allowed A configuration a
  // (its A-set), and a slice is up to 64 consecutive rows of one group. Terms
  // that move B are stored once per group as entries (source group's first
  // row, term, sign from b); a lane applies an entry to its own
  // a, either unchanged (identity: the source rows are contiguous) or through
  // a map a -> source index. Terms acting on A alone form a small matrix per
  // A-set (4-byte slots: source index | term | sign).
  int kind = 0;                        // 0: SELL slices of 64 rows, 1: product layout
  const int4 *pslice = nullptr;        // per slice: {first row, rows (<= 64), group, A-set | chunk << 16}
  const uint2 *pent_ptr = nullptr;     // per group: {first entry, first mapped entry} (ngroups + 1)
  const uint4 *pent = nullptr;         // {source group's first row, term | sign << 31, map offset or ~0, 0}
  const int32_t *amap = nullptr;       // per map and destination index: source index (bit 30: sign) or -1
  const int4 *aset = nullptr;          // per A-set: {rows m, A-matrix ELL width k, ELL offset, 0}
  const uint32_t *hA = nullptr;        // A-matrix ELL, column-major per A-set: src | term << 16 | sign << 31
};

// Host side of a product layout, kept with the shared index.
struct ProdHost {
  std::vector<int> diag_terms;         // terms with F == 0, in term order
  int64_t ngroups = 0, nsets = 0, nmaps = 0, nentries = 0, nmapped = 0, hA_slots = 0;
  double load_cost = 0.0;              // estimated 64-lane 16-byte vector loads per apply
};

// Device arrays of an assembled index stream (everything but the term tables
// and the diagonal), shared by the bindings made from it with
// qs_hop_create_like and freed when the last of them goes.
struct HopIndex {
  std::vector<void *> buf;
  std::shared_ptr<ProdHost> prod;  // product layout only
  ~HopIndex();
};

struct qs_hop {
  const qs_op *op = nullptr;
  const qs_basis *rows = nullptr;
  const qs_basis *cols = nullptr;
  bool same = false;
  int assembled = 0;
  int64_t nnz_eff = 0;             // (row, term) pairs with a source in cols
  int64_t nnz_stored = 0;          // stored index entries: ELL slots (incl. padding) + block headers
  int64_t nnz_offdiag = 0;         // off-diagonal entries
  int64_t stream_bytes = 0;        // bytes of the assembled index stream and per-row data
  int64_t nblocks = 0;             // shift blocks
  int64_t block_hits = 0;          // off-diagonal entries held by shift blocks
  int64_t ell_slots = 0;           // ELL slots incl. padding
  int shift_min = 0;               // lanes a group needs to become a shift block
  int64_t nslices = 0;
  SellD view;                      // device view (pointers below)
  uint64_t *d_blk_ptr = nullptr;
  uint4 *d_blk = nullptr;
  uint64_t *d_slice_ptr = nullptr;
  uint32_t *d_slice_len = nullptr;
  uint16_t *d_rowcnt = nullptr;
  uint32_t *d_col = nullptr;
  uint16_t *d_tid = nullptr;
  double *d_diag = nullptr;
  std::shared_ptr<HopIndex> index;  // owner of the d_* arrays above except d_diag, once assembled
  // product layout: canonical index of each row (the index owns it)
  const int32_t *d_inv = nullptr;
};

// ---- error helpers --------------------------------------------------------
qs_status qs_fail(qs_ctx *ctx, qs_status st, const char *fmt, ...);
qs_status qs_hip_fail(qs_ctx *ctx, hipError_t e, const char *what, const char *file, int line);

#define QS_HIP(ctx, call)                                                       \
  do {                                                                          \
    hipError_t _e = (call);                                                     \
    if (_e != hipSuccess) return qs_hip_fail((ctx), _e, #call, __FILE__, __LINE__); \
  } while (0)

#define QS_TRY(call)                                                            \
  do {                                                                          \
    qs_status _s = (call);                                                      \
    if (_s != QS_OK) return _s;                                                 \
  } while (0)

qs_status qs_malloc(qs_ctx *ctx, void **p, size_t bytes);
// Host <-> device copies staged through the ctx's own pinned buffer, ordered
// on ctx->stream; both return after the copy (and the stream's earlier work)
// has completed. Copies from pageable memory go through the runtime's shared
// staging path, where one thread's small copy waits behind another thread's
// large one; with a pinned buffer per ctx, threads do not wait on each other.
qs_status qs_h2d(qs_ctx *ctx, void *dst, const void *src, size_t bytes);
qs_status qs_d2h(qs_ctx *ctx, void *dst, const void *src, size_t bytes);
// Return a qs_malloc block to the library's device memory cache (null is a
// no-op). The block must no longer be in use by queued work.
void qs_free(void *p);
qs_status qs_scratch(qs_ctx *ctx, size_t bytes, void **p);

// Workgroups of QS_BLOCK (256) threads of `kernel` with `shm` bytes of dynamic
// LDS that the whole device keeps resident at once. Grid-stride kernels are
// launched with at most this many, so no workgroup waits for a slot while the
// others already drain (which would leave its share of the work as a tail).
int64_t qs_resident_blocks(qs_ctx *ctx, const void *kernel, size_t shm);

// ---- internal entry points across translation units ----------------------
// count_only: fill rt.total only (saturated at INT64_MAX), with no size limit
// and no tables; otherwise sectors of 2^32-1 or more determinants are refused.
qs_status qs_build_rank_table(int nf, int nres, const uint8_t *masks, const int *lo,
                              const int *hi, RankTable &rt, std::string &err, bool count_only);
qs_status qs_dot_internal(qs_ctx *ctx, int64_t n, const double *a, const double *b,
                          double *re, double *im);

// Index in b of n determinants in device memory (-1 when absent), into d_idx
// (device), enqueued on ctx->stream (qs_core.hip).
qs_status qs_basis_index_dev(qs_ctx *ctx, const qs_basis *b, const uint64_t *d_words, int64_t n, int64_t *d_idx);
// Pre-summed diagonal of n rows over the terms d_which (device, in term order)
// into diag, enqueued on ctx->stream (qs_apply.hip).
qs_status qs_diag_rows(qs_ctx *ctx, const qs_op *op, const uint64_t *rows, int64_t n, const int *d_which, int nwhich,
                       void *diag, int real_diag);
// Product layout (qs_product.hip): a binding sharing `like`'s product index
// with op's own values; the vector order conversions; apply in the binding's
// row order.
qs_status qs_prod_create_like(qs_ctx *ctx, const qs_op *op, const qs_hop *like, qs_hop **out);
qs_status qs_prod_to_rows(qs_ctx *ctx, const qs_hop *h, int k, const double2 *x, int64_t ldvx, double2 *xp,
                          int64_t ldvp);
qs_status qs_prod_from_rows(qs_ctx *ctx, const qs_hop *h, int k, const double2 *yp, int64_t ldvp, double2 *y,
                            int64_t ldvy);
qs_status qs_prod_apply_rows(qs_ctx *ctx, const qs_hop *h, int k, const double2 *x, int64_t ldvx, double2 *y,
                             int64_t ldvy);

// Householder reduction of the n x n Hermitian A (device, column-major, lower
// triangle as zhetrd uses it) in one cooperative launch, with Y <- Q^+ Y for
// nv vectors (qs_tridiag.hip); QS_EUNSUPPORTED, nothing done, when n is too
// large or the device cannot hold the grid.
qs_status qs_hetrd_fused(qs_ctx *ctx, int n, double2 *A, double *d, double *e, double2 *tau, int nv, double2 *Y);

// Host-side dense Hermitian eigen-solver (column-major, complex interleaved).
// A: n*n complex (overwritten), w: n eigenvalues ascending, Z: n*n eigenvectors
// (column j = eigenvector j).
void qs_heev(int n, const double *A, double *w, double *Z);
// all eigenvalues ascending, eigenvectors of the nvec lowest (Z: n x nvec
// column-major); false if the QL iteration does not converge
bool qs_heev_lowest(int n, const double *A, double *w, double *Z, int nvec);

// ---- device helpers -------------------------------------------------------
__device__ __forceinline__ uint32_t qs_chunk_byte1(uint64_t w, int c) {
  return (uint32_t)(w >> (56 - 8 * c)) & 0xFFu;
}
