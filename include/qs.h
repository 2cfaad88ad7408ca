/*
 * qs.h — C ABI of libqs, the MI355X engine for the second-quantization hot path:
 * operator x wavefunction on Slater-determinant bitstrings (H*psi), Lanczos
 * tridiagonalization, block-Lanczos, lowest eigenstates and the
 * continued-fraction Green's function.
 *
 * Conventions shared by every entry point
 *  - Determinants of NF <= 128 fermion modes are packed into NW = 1 (NF <= 64)
 *    or NW = 2 (NF <= 128) uint64 words, word 0 most significant. Mode i sits at
 *    bit (64*NW - 1 - i) of the 64*NW-bit integer, so numeric order equals the
 *    memcmp order of the reference's byte layout (mode i = bit 128>>(i&7) of
 *    byte i>>3, Core/OperatorWaveFunction.cpp:10-12).
 *  - A ladder term of length l is l uint16 codes o[0..l-1]; bit 15 set = create,
 *    low 15 bits = mode. o[0] is applied first, i.e. it is the rightmost written
 *    factor (Core/StructDef.cpp:46-47, Core/OperatorWaveFunction.cpp:98-116).
 *  - Vectors are complex FP64, interleaved (re, im), one entry per basis
 *    determinant in the basis' order, resident in device memory (HBM).
 *  - Every function returns a qs_status; on failure qs_last_error() holds a
 *    message. No function calls exit().
 *  - One qs_ctx per host thread. A ctx owns one HIP stream; all device work of
 *    a call is ordered on it, and calls return after that work has completed
 *    unless stated otherwise.
 */
#ifndef QS_H
#define QS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  QS_OK = 0,
  QS_ENOMEM_DEVICE = 1, /* hipMalloc failed (cf. reference status 1 = allocation) */
  QS_ENOMEM_HOST = 2,
  QS_EINVAL = 3,
  QS_EHIP = 4,
  QS_ENOTCONVERGED = 5,
  QS_EUNSUPPORTED = 6
} qs_status;

typedef struct qs_ctx qs_ctx;     /* device, stream, scratch, last error */
typedef struct qs_basis qs_basis; /* sorted determinant set on the device (+ rank tables) */
typedef struct qs_op qs_op;       /* compiled ladder-term table on the device */
typedef struct qs_hop qs_hop;     /* an operator bound to (row basis, column basis) */

/* ---- context ------------------------------------------------------------ */
qs_status qs_ctx_create(int device, qs_ctx **ctx);
qs_status qs_ctx_destroy(qs_ctx *ctx);
const char *qs_last_error(const qs_ctx *ctx);
/* Device memory helpers so C callers need no HIP headers. All device memory
 * of the library (these, and what bases, operators, batches and solvers hold)
 * is carved from chunks of device memory that stay with the process (at least
 * 4 GiB each; a larger block gets a chunk of its own size): the driver clears
 * memory when it hands it out, tens of milliseconds per GB, so a stream of
 * requests that returned and re-took its multi-GB blocks would spend seconds
 * in the driver. A freed block is merged with its free neighbours and reused
 * for later blocks of any size; allocating and freeing one large block
 * reserves that much for the library. Wholly free chunks go back to the
 * driver when a new chunk would leave less than 8 GiB of the device free for
 * other allocators, and by qs_trim_memory. Free or destroy only after the work
 * that uses the memory has completed (qs_sync for an async context).
 * QS_ALLOC_TRACE=1 in the environment reports every chunk taken from or
 * returned to the driver on stderr. */
qs_status qs_dev_alloc(qs_ctx *ctx, size_t bytes, void **ptr);
qs_status qs_dev_free(qs_ctx *ctx, void *ptr);
/* Return wholly free chunks of ctx's device to the driver, largest first,
 * until at most keep_bytes stay free (0: every wholly free chunk). The
 * library also gives chunks back by itself when an allocation would otherwise
 * fail. */
qs_status qs_trim_memory(qs_ctx *ctx, size_t keep_bytes);
/* Bytes in live blocks and free bytes held in chunks, on ctx's device. */
qs_status qs_memory_stats(qs_ctx *ctx, size_t *live_bytes, size_t *cached_bytes);
qs_status qs_memcpy_h2d(qs_ctx *ctx, void *dst, const void *src, size_t bytes);
qs_status qs_memcpy_d2h(qs_ctx *ctx, void *dst, const void *src, size_t bytes);
qs_status qs_sync(qs_ctx *ctx);
/* The ctx's HIP stream (a hipStream_t), for callers that order their own work
 * or record timing events on it. */
void *qs_ctx_stream(qs_ctx *ctx);
/* async != 0: entry points whose results stay on the device (qs_hop_apply,
 * qs_hop_apply_multi, qs_apply) return after enqueueing their kernels instead
 * of waiting for them; the caller orders later work on qs_ctx_stream() or
 * calls qs_sync(). Default 0. */
qs_status qs_ctx_set_async(qs_ctx *ctx, int async);
/* Tuning parameters of a ctx (no counterpart in the reference, whose
 * operator application has no assembled form).
 * QS_PARAM_SHIFT_MIN (default 56, range 1..65): in operators assembled after
 * the call, a term's hits inside a 64-row slice whose columns are a common
 * shift of the row positions are stored as one block header when at least
 * this many rows share the shift; 65 stores every entry individually.
 * QS_PARAM_ELL_PACK (default 1): 1 stores an individual entry as one uint32
 * (column and term id) when both fit, 0 always as uint32 column + uint16 term.
 * QS_PARAM_BATCH_GRAPH (default 1): 1 runs qs_batch_run's steps as one HIP
 * graph launch (a chain of kernel nodes per batch and step count, kept across
 * slot changes); 0 launches the kernels one by one. Results are identical.
 * QS_PARAM_TERM_LDS (default 24576, range 0..65536): assembled H*psi kernels
 * stage an operator's term tables (value and parity mask per term) in LDS when
 * they take at most this many bytes, and read them from global memory
 * otherwise. Larger tables in LDS leave fewer workgroups per CU.
 * QS_PARAM_ELL_ORDER (default 1): the order of a row's individually stored
 * entries in operators assembled after the call; 0 keeps the term order, 1
 * sorts them by column, so that the lanes of a slice gather from nearby
 * columns at each entry position (2% faster H*psi on the 2,179,235-row
 * impurity). Results agree to rounding (the sum runs in another order).
 * QS_PARAM_ALLOC_TOP (default 0): 1 places this context's later device
 * allocations at the top of the library's free device memory, for data that
 * stays long (a cache of assembled operators): the free space below then
 * stays in large pieces for the short-lived blocks placed best-fit.
 * QS_PARAM_HETRD (default 1): the Householder reduction of the dense method
 * (qs_dense_poles, qs_dense_lowest); 1 runs it as one cooperative launch
 * with the matrix held in registers across the CUs (n <= 1536 on the MI355X),
 * else rocSOLVER's zhetrd; 0 always uses zhetrd. Both give LAPACK's zhetd2
 * reduction; results agree to rounding.
 * QS_PARAM_TRIDIAG_DC (default 256, >= 1): the order from which the dense
 * method solves its tridiagonal eigenproblem on the GPU by rocSOLVER's divide
 * and conquer (dstedc); below it the host QL runs (faster for small orders).
 * All eigenvectors (qs_dense_poles with vectors = 1) always take dstedc.
 * QS_PARAM_WAIT_SPINS (default 2^22, 0 to 2^40): how many times a wait inside
 * the register-resident reduction polls for data another workgroup has not
 * written yet before the launch gives up (every workgroup then stops, and the
 * call returns QS_EHIP naming the wait); a bound, not a timing knob. */
enum { QS_PARAM_SHIFT_MIN = 1, QS_PARAM_ELL_PACK = 2, QS_PARAM_BATCH_GRAPH = 3, QS_PARAM_TERM_LDS = 4,
       QS_PARAM_ELL_ORDER = 5, QS_PARAM_ALLOC_TOP = 6, QS_PARAM_HETRD = 7, QS_PARAM_TRIDIAG_DC = 8,
       QS_PARAM_WAIT_SPINS = 9 };
qs_status qs_ctx_set_param(qs_ctx *ctx, int param, int64_t value);
qs_status qs_ctx_get_param(qs_ctx *ctx, int param, int64_t *value);

/* ---- basis (restricted-occupation determinant sets) --------------------- */
/* Enumerate every determinant of nf modes whose occupation inside each mask r
 * lies in [lo[r], hi[r]] (the test of QDetNotRestricted,
 * Core/OperatorWaveFunction.cpp:2260-2286, restriction tables as parsed by
 * luaL_setarray_DeterminatRestrictionType, TableToArray.cpp:3005-3068).
 * masks: nres rows of ceil(nf/8) bytes in the reference byte layout.
 * The result is sorted ascending and carries exact rank tables, so lookups are
 * O(ceil(nf/8)) table steps. Replaces the reference's hash-set basis
 * (Core/Hash.cpp, Core/ComplexWaveFunctionArithmetic.cpp:64-353). */
qs_status qs_basis_enumerate(qs_ctx *ctx, int nf, int nres, const uint8_t *masks,
                             const int *lo, const int *hi, qs_basis **out);
/* Count only (no device allocation of the determinant list, no size limit);
 * counts above INT64_MAX saturate at INT64_MAX. qs_basis_enumerate refuses
 * sectors of 2^32 - 1 or more determinants. */
qs_status qs_basis_count(qs_ctx *ctx, int nf, int nres, const uint8_t *masks,
                         const int *lo, const int *hi, int64_t *n);
/* Build a basis from an explicit list of n determinants (host, NW words each,
 * any order, duplicates removed). Lookups use binary search. */
qs_status qs_basis_from_words(qs_ctx *ctx, int nf, const uint64_t *words, int64_t n,
                              qs_basis **out);
/* Closure of a determinant set under an operator: the smallest set containing
 * `seed` and every target t|d> != 0 of op's terms (a term counts when it acts
 * on the determinant, whatever its amplitude), optionally filtered by at most
 * 16 restrictions (nres may be 0), iterated until nothing new appears. When
 * the set passes max_dim the call returns QS_ENOTCONVERGED and *out = NULL.
 * This is the reachable space the reference grows by repeated H*psi
 * (Core/BlockLanczos.cpp:5849-5858, LuaInterface/ssp.cpp:2265-2316). Built on
 * the GPU; the result is sorted, lookups use binary search. */
qs_status qs_basis_closure(qs_ctx *ctx, const qs_op *op, const qs_basis *seed, int nres,
                           const uint8_t *masks, const int *lo, const int *hi,
                           int64_t max_dim, qs_basis **out);
/* Image of a determinant set under an operator: every target t|d> != 0 of
 * op's terms over the determinants d of `seed` (a term counts when it acts on
 * d, whatever its amplitude), inside at most 16 restrictions; sorted, lookups
 * by binary search. This is the output basis of the push-form OperatorPsi /
 * OperatorPsiRestricted (Core/OperatorWaveFunction.cpp:1750-2258, 2288-2794):
 * the determinants O|psi> reaches from psi's support, with no enumeration of
 * the whole particle sector. QS_ENOTCONVERGED and *out = NULL when the image
 * holds more than max_dim determinants. */
qs_status qs_basis_image(qs_ctx *ctx, const qs_op *op, const qs_basis *seed, int nres,
                         const uint8_t *masks, const int *lo, const int *hi, int64_t max_dim,
                         qs_basis **out);
int64_t qs_basis_size(const qs_basis *b);
int qs_basis_nf(const qs_basis *b);
int qs_basis_words(const qs_basis *b); /* NW: 1 or 2 */
/* Copy the sorted determinant words to host (n*NW uint64). */
qs_status qs_basis_export(qs_ctx *ctx, const qs_basis *b, uint64_t *words);
/* Device pointer to the sorted determinant words (n*NW uint64). */
const uint64_t *qs_basis_device_words(const qs_basis *b);
/* Index of each of n host determinants in b (-1 when absent). */
qs_status qs_basis_index(qs_ctx *ctx, const qs_basis *b, const uint64_t *words, int64_t n,
                         int64_t *idx);
qs_status qs_basis_destroy(qs_basis *b);

/* ---- operators ---------------------------------------------------------- */
/* nterms ladder strings: len[t] codes each, concatenated in `codes`;
 * values: 2*nterms doubles (re, im). Terms that vanish identically (e.g. c+_i c+_i)
 * are dropped. Fermions only (the reference rejects NB>0 wavefunctions,
 * LuaInterface/ssp.cpp:950-953). Replaces OperatorType / OperatorFixedLengthType
 * (Core/StructDef.cpp:223-249) as input to OperatorPsi (Core/OperatorWaveFunction.cpp:5598). */
qs_status qs_op_create(qs_ctx *ctx, int nf, int nterms, const int *len, const uint16_t *codes,
                       const double *values, qs_op **out);
int qs_op_nterms(const qs_op *op);
qs_status qs_op_destroy(qs_op *op);

/* ---- operator on a basis: apply ---------------------------------------- */
enum { QS_HOP_MATFREE = 0, QS_HOP_ASSEMBLED = 1, QS_HOP_AUTO = 2 };
/* Bind op to rows = `rows` basis and columns = `cols` basis. ASSEMBLED builds
 * the connection stream once and reuses it on every apply: per 64-row slice,
 * shift blocks (a term whose rows read consecutive columns, one header each)
 * and the remaining entries (column index + term id each); MATFREE recomputes
 * targets on the fly; AUTO assembles when the stream fits in
 * `mem_budget_bytes` (0 = 50% of free HBM). */
qs_status qs_hop_create(qs_ctx *ctx, const qs_op *op, const qs_basis *rows, const qs_basis *cols,
                        int mode, size_t mem_budget_bytes, qs_hop **out);
/* Bind op to the bases of `like`, an assembled binding whose operator compiles
 * to the same terms (masks and order; values may differ), sharing like's
 * connection stream: only the term values and the pre-summed diagonal are
 * op's own, so nothing is enumerated or assembled. The shared stream lives
 * until the last binding that uses it is destroyed; the bases must outlive
 * every such binding. QS_EINVAL with a message when the terms differ. No
 * counterpart in the reference, which has no assembled operator: for callers
 * that apply one operator family with many parameter sets (a server's
 * requests of one impurity shape). */
qs_status qs_hop_create_like(qs_ctx *ctx, const qs_op *op, const qs_hop *like, qs_hop **out);
/* Bind op to a restricted sector `basis` (rows == cols) in the product
 * layout: the modes are split into group A (amask: NF bits, mode i at bit
 * 7-(i&7) of byte i>>3, as restriction masks) and group B, every restriction
 * of the basis must lie inside A, inside B or cover all modes, and rows are
 * grouped by their B configuration. Terms that move B are stored once per
 * group and applied to all of a group's rows with contiguous loads; terms on
 * A alone form small per-A-set matrices. Same results as an ASSEMBLED binding
 * (OperatorPsiConserveBasis, Core/OperatorWaveFunction.cpp:245) up to the
 * summation order; vectors passed to qs_hop_apply, qs_lanczos and the block
 * and ground-state entry points stay in canonical order. QS_EUNSUPPORTED with
 * a message when the sector does not factor over the split. No counterpart in
 * the reference (its hash-set H*psi has no layout); for impurity models whose
 * bath moves dominate. qs_batch_add does not take such bindings. */
qs_status qs_hop_create_product(qs_ctx *ctx, const qs_op *op, const qs_basis *basis, const uint8_t *amask,
                                qs_hop **out);
/* Shape of a product-layout binding: info[0..7] = groups, slices, A-sets,
 * entries, mapped entries, 0 (reserved), A-matrix slots, maps; load_cost
 * (may be NULL) = the estimated 64-lane 16-byte vector loads of one apply,
 * for choosing the mode split. QS_EINVAL for other bindings. */
qs_status qs_hop_product_info(const qs_hop *h, int64_t *info, double *load_cost);
/* The canonical index of every row of h in the order its kernels keep
 * vectors (n int64, host): the identity except for product-layout bindings. */
qs_status qs_hop_row_order(qs_ctx *ctx, const qs_hop *h, int64_t *canon);
/* y = O x with x and y in h's own row order (qs_hop_row_order): the apply
 * kernel alone, without order conversions. */
qs_status qs_hop_apply_rows(qs_ctx *ctx, const qs_hop *h, const double *x, double *y);
/* Number of (row, term) pairs with nonzero phase whose source lies in cols,
 * the stored off-diagonal stream length (incl. slice padding; diagonal terms
 * are pre-summed), its bytes, and the off-diagonal entries without padding.
 * Any pointer may be NULL. */
qs_status qs_hop_stats(const qs_hop *h, int64_t *nnz_eff, int64_t *nnz_stored,
                       int64_t *stream_bytes, int *assembled, int64_t *nnz_offdiag);
/* Layout of an assembled binding: shift blocks, the entries they hold, the
 * remaining per-entry slots (incl. padding), whether an entry is one packed
 * uint32, and the shift threshold used. Any pointer may be NULL. */
qs_status qs_hop_layout(const qs_hop *h, int64_t *nblocks, int64_t *block_hits, int64_t *ell_slots,
                        int *packed, int *shift_min);
/* The diagonal <d_i|O|d_i> of an assembled binding with rows == cols, as n
 * complex numbers into device memory d (no counterpart in the reference; used
 * to weight Lanczos start vectors towards low-energy determinants). */
qs_status qs_hop_diagonal(qs_ctx *ctx, const qs_hop *h, double *d);
/* y = O x (pull form: y_i = sum_t V_t <d_i| t |d_j> x_j). x over cols, y over
 * rows, device pointers. Equals OperatorPsiRestricted (Core/OperatorWaveFunction.cpp:2288)
 * with the row basis as restriction set, and OperatorPsiConserveBasis (:245) when
 * rows == cols. */
qs_status qs_hop_apply(qs_ctx *ctx, const qs_hop *h, const double *x, double *y);
/* Same for k vectors stored back to back (vector v at x + 2*v*ncols). */
qs_status qs_hop_apply_multi(qs_ctx *ctx, const qs_hop *h, int k, const double *x, double *y);
qs_status qs_hop_destroy(qs_hop *h);

/* One-shot apply through a temporary matrix-free binding. */
qs_status qs_apply(qs_ctx *ctx, const qs_op *op, const qs_basis *rows, const qs_basis *cols,
                   const double *x, double *y);
/* <x|O|x> over one basis (PsiOperatorPsi, Core/OperatorWaveFunction.cpp:5680). */
qs_status qs_expect(qs_ctx *ctx, const qs_op *op, const qs_basis *b, const double *x,
                    double *re, double *im);

/* <x|O|x> with no allocation at all: no counterpart in the reference. For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The same value as
 * qs_expect, written to out (device, 2 doubles: re, im) instead of the host.
 * O x and the partial sums live in a static device array of 64 MiB that every
 * GPU the library runs on holds until the process ends, so the basis may hold
 * at most 4,193,280 determinants (larger bases are refused with QS_EINVAL;
 * use qs_expect). Calls that overlap on one GPU share that array and race:
 * serialize them (one stream, or one call at a time). With an async ctx the
 * call returns after enqueueing its kernels. */
qs_status qs_expect_noalloc(qs_ctx *ctx, const qs_op *op, const qs_basis *b, const double *x, double *out);

/* ---- vector helpers (device pointers, n complex entries) ---------------- */
/* <a|b> with a conjugated (ComplexWaveFunctionDotProductSameBasis,
 * Core/ComplexWaveFunctionArithmetic.cpp:961). Fixed-order, deterministic. */
qs_status qs_dot(qs_ctx *ctx, int64_t n, const double *a, const double *b, double *re,
                 double *im);
qs_status qs_norm(qs_ctx *ctx, int64_t n, const double *a, double *nrm);

/* ---- Lanczos tridiagonalization (spectra) ------------------------------- */
/* Scalar Lanczos on h (rows == cols) from v0 (device, not modified):
 * a[0]=0, b[0]=|v0|, a[i]=Re<Hv_{i-1}|v_{i-1}>, b[i]=|r_i|, no
 * reorthogonalization — the recurrence of LanczosTriDiagonalizeCC
 * (Core/Lanczos.cpp:2169-2264) on a fixed basis, epsilon = 0.
 * a_out: nsteps+1 doubles, b_out: nsteps doubles (host). */
qs_status qs_lanczos(qs_ctx *ctx, const qs_hop *h, const double *v0, int nsteps, double *a_out,
                     double *b_out);
/* Lanczos that keeps its Krylov basis (CreateSpectra with KrylovBasis = true):
 * LanczosTriDiagonalizeKrylovCC (Core/Lanczos.cpp:3705-3854). Each new vector
 * is orthogonalized twice against every kept vector, normalized, and set to
 * zero when its squared overlaps with the kept vectors still exceed eps (the
 * Krylov space is exhausted); a[nsteps] = <v|H|v> of the last vector.
 * a_out: nsteps+1, b_out: nsteps doubles (host); V_out: device, nsteps vectors
 * of n complex back to back (vector i at V_out + 2*i*n), v_0 = v0/|v0|. */
qs_status qs_lanczos_krylov(qs_ctx *ctx, const qs_hop *h, const double *v0, int nsteps, double eps,
                            double *a_out, double *b_out, double *V_out);
/* Block Lanczos for k start vectors V0 (k vectors back to back, device):
 * block recurrence of LanczosBlockTriDiagonalizeCC (Core/BlockLanczos.cpp:2131-2371):
 * Loewdin-orthonormalised start block, B0 = <v|V0>, per step A_k = V_k^+ H V_k,
 * local re-orthogonalization against the two previous blocks (two classical
 * Gram-Schmidt rounds against both together), singular values below sv_cut
 * dropped. A_out: (nsteps+1)*k*k complex, B_out: nsteps*k*k complex
 * (row-major blocks, host); *kept receives the block size of each step
 * (nsteps+1 ints, may be NULL). */
qs_status qs_block_lanczos(qs_ctx *ctx, const qs_hop *h, int k, const double *V0, int nsteps,
                           double sv_cut, double *A_out, double *B_out, int *kept);

/* ---- batched Lanczos (many problems in lockstep) ------------------------ */
/* No single counterpart in the reference, which runs one (T, psi) Lanczos at a
 * time (LuaInterface/ssp.cpp:2598-2639 loops over LanczosTriDiagonalize). A
 * batch holds up to max_slots independent recurrences; every qs_batch_run step
 * advances all of them with one set of launches. A slot's coefficients are
 * bit-identical whether it runs alone or with others. */
typedef struct qs_batch qs_batch;
qs_status qs_batch_create(qs_ctx *ctx, int max_slots, qs_batch **out);
/* Start a recurrence of `depth` steps on an assembled h (rows == cols; the
 * sliced ELL or the product layout, whose slot keeps its vectors in the
 * layout's row order while v0, Ritz vectors and y stay canonical) from
 * device vector v0 (copied); max_depth (>= depth) sizes the coefficient arrays
 * so qs_batch_set_depth can extend the run. ycoef (host, depth doubles, or
 * NULL): also build y = sum_i ycoef[i] v_i, the Ritz vector of a previous pass
 * over the same start. */
qs_status qs_batch_add(qs_batch *b, const qs_hop *h, const double *v0, int depth, int max_depth,
                       const double *ycoef, int *slot);
/* qs_batch_add with flags. QS_BATCH_KEEP_BASIS: the slot also keeps its
 * Lanczos vectors on the device ((max_depth + 1) * n complex), so that
 * qs_batch_ritz forms a Ritz vector without a second pass. */
enum { QS_BATCH_KEEP_BASIS = 1 };
qs_status qs_batch_add_ex(qs_batch *b, const qs_hop *h, const double *v0, int depth, int max_depth, int flags,
                          const double *ycoef, int *slot);
/* y = sum_{i<m} z[i] v_i over the kept normalized Lanczos vectors v_i of a
 * QS_BATCH_KEEP_BASIS slot (m <= completed steps; z host, y device, n complex).
 * The same sum, in the same order, as a second pass with ycoef = z. */
qs_status qs_batch_ritz(qs_batch *b, int slot, int m, const double *z, double *y);
/* Change the depth of a coefficient-only slot (<= its max_depth); a finished
 * slot resumes from its last step when the depth grows. */
qs_status qs_batch_set_depth(qs_batch *b, int slot, int depth);
/* Advance every active slot by `steps` Lanczos steps (slots stop at their depth). */
qs_status qs_batch_run(qs_batch *b, int steps);
/* Completed steps of a slot; a/b (host, depth+1 doubles each, may be NULL)
 * receive a[0..depth] and b[0..depth] (b[depth] is the norm after the last
 * step); y (device, n complex, may be NULL) receives the accumulated vector. */
qs_status qs_batch_poll(qs_batch *b, int slot, int *step, double *a, double *bb, double *y);
qs_status qs_batch_remove(qs_batch *b, int slot);
int qs_batch_active(const qs_batch *b);
/* Coefficients of n slots in one transfer: steps[n] completed steps, a and b
 * (host, n x ld, row i = slots[i]) receive a[0..ld) and b[0..ld) of each slot
 * (zero past its max_depth). */
qs_status qs_batch_coeffs(qs_batch *b, int n, const int *slots, int ld, int *steps, double *a, double *bb);
/* qs_batch_set_depth for n slots with one launch. */
qs_status qs_batch_set_depths(qs_batch *b, int n, const int *slots, const int *depths);
/* Completed steps of every slot in one transfer: steps[max_slots], -1 for free slots. */
qs_status qs_batch_steps(qs_batch *b, int *steps);
qs_status qs_batch_destroy(qs_batch *b);

/* ---- lowest eigenstates (Eigensystem) ----------------------------------- */
/* nstates lowest eigenpairs of Hermitian h (rows == cols) by thick-restarted
 * block Lanczos with full re-orthogonalization; block size = max(nstates, 1).
 * Converged when every wanted state has |H y - E y|^2 <= tol * max(1, sum E^2)/nstates
 * (the variance test of BlockGroundState, Core/BlockLanczos.cpp:5869-5883).
 * start: nstates device vectors or NULL (seeded pseudo-random start).
 * E_out: nstates doubles (host, ascending); vecs: nstates device vectors (out). */
qs_status qs_ground_states(qs_ctx *ctx, const qs_hop *h, int nstates, double tol, int max_krylov,
                           int max_restarts, uint64_t seed, const double *start, double *E_out,
                           double *vecs, double *variance_out, int *iterations_out);
/* qs_ground_states with the Krylov schedule spelled out: the basis holds at
 * most nstates + krylov_start vectors before the first thick restart and grows
 * by krylov_step + nstates per restart while that stays below max_krylov — the
 * NKrylovStart / NKrylovStep / NKrylovMax settings (defaults 100 / 10 / 500,
 * Core/StructDef.cpp:298-300) as BlockLanczosGroundStateConserveBasisKrylov
 * uses them (Core/BlockLanczos.cpp:5130, 5311-5325). qs_ground_states is this
 * with krylov_start 100, krylov_step 10 and flags 0. */
qs_status qs_ground_states_ex(qs_ctx *ctx, const qs_hop *h, int nstates, double tol, int krylov_start,
                              int krylov_step, int max_krylov, int max_restarts, uint64_t seed, int flags,
                              const double *start, double *E_out, double *vecs, double *variance_out,
                              int *iterations_out);
/* flags of qs_ground_states_ex. QS_GS_CYCLE_END: test convergence only when the
 * basis reaches its restart size, as the reference's Krylov loop does (its
 * deflation test sees the Ritz residual block once per restart cycle,
 * Core/BlockLanczos.cpp:5176-5215). Without it the Ritz pairs are also tested
 * at intermediate basis sizes and the solver may stop earlier; that stops
 * within the same variance bound but resolves states that are much closer
 * than the bound's residual (a field-split multiplet) less sharply. */
#define QS_GS_CYCLE_END 1

/* Tensor Green's function from qs_block_lanczos output, evaluated on the GPU:
 * G(w) = B0^* G_1 B0^T, G_j = (z - A_j - B_j G_{j+1} B_j^+)^-1 bottom-up,
 * z = w + E0 + i Gamma/2, w on the same ne+1 point grid as qs_cf_spectrum —
 * BlockTriDiagonalMatrixToSpectraWithBroadening (BasicMath_BlockTypeConversion.cpp:1636-1726).
 * A: (n+1)*k*k, B: n*k*k complex, kept: n+1 ints (as written by
 * qs_block_lanczos; k <= 16); out: (ne+1)*k*k complex, out[i][r][c] =
 * <X_r|(z-H)^-1|X_c> for the start vectors X. Host arrays. */
qs_status qs_block_cf_spectrum(qs_ctx *ctx, int k, int n, const double *A, const double *B,
                               const int *kept, double E0, double emin, double emax, int ne,
                               double gamma, double *out);

/* Dense matrix of a binding on the device: M[c*nrows + r] = <r|O|c> (complex,
 * column-major, nrows x ncols), the entries of every term added in term order
 * — OperatorToDeterminantMatrix (Core/MatrixOperator.cpp:1149-1226). M: device,
 * nrows*ncols complex. */
qs_status qs_hop_dense(qs_ctx *ctx, const qs_hop *h, double *M);

/* Eigen-decomposition of a dense Hermitian n x n matrix A (device, complex,
 * column-major; its lower triangle is read) and the projections of nv vectors
 * X (device, vector s at X + 2*s*n) on the eigenvectors z_k:
 * E[k] (host, ascending), P[k*nv + s] = <z_k|x_s> (host, complex).
 * |P[k][s]|^2 are the pole weights of HermitianCompactMatrixToListOfPoles
 * (Core/BasicMath/BasicMath_CompactMatrixVector.cpp:3990-4026), and
 * conj(P[k][r]) P[k][c] the pole blocks of ...ToBlockListOfPoles (:4033-4113).
 * Method: Householder reduction A = Q T Q^+ on the GPU with Q^+ X formed
 * there (QS_PARAM_HETRD); the tridiagonal eigenvectors Z of T by divide and
 * conquer on the GPU and P = Z^T Q^+ X (from order QS_PARAM_TRIDIAG_DC on), or
 * the tridiagonal QL on the host with its rotations applied to Q^+ X.
 * vectors = 1: A is overwritten by the eigenvectors Q Z (column k = z_k).
 * vectors = 0: eigenvalues and projections only, A is overwritten by the
 * reduction.
 * Within a degenerate eigenvalue the eigenvector basis, hence the individual
 * P, depends on the method; sums over the degenerate set do not.
 * QS_ENOTCONVERGED when the solver does not converge. */
qs_status qs_dense_poles(qs_ctx *ctx, int n, double *A, int nv, const double *X, double *E, double *P,
                         int vectors);

/* All eigenvalues E (host, ascending) of a dense Hermitian n x n matrix A
 * (device, complex, column-major, lower triangle read, overwritten) and the
 * eigenvectors of the k lowest in V (device, n x k complex, column j = z_j):
 * Householder reduction on the GPU, the tridiagonal eigenvectors by divide and
 * conquer on the GPU (from order QS_PARAM_TRIDIAG_DC on) or by the host QL
 * with its rotations replayed on k unit vectors, back-transform on the GPU. The
 * dense ground states of BlockLanczosGroundStateConserveBasisDense
 * (Core/BlockLanczos.cpp:370-504). */
qs_status qs_dense_lowest(qs_ctx *ctx, int n, double *A, int k, double *E, double *V);

/* Spectra from poles E[0..npole) with projections P (npole*nv complex, as
 * written by qs_dense_poles), on the GPU; host arrays.
 * tensor = 0: out[s*(ne+1) + i] = sum_k |P[k][s]|^2 / (w_i + E0[s] - E_k + i gamma/2),
 *   w_i = emin + i (emax-emin)/ne — ListOfPolesToSpectrum (Core/Spectral.cpp:1516-1591).
 * tensor = 1: out[(i*nv + r)*nv + c] = sum_k conj(P[k][r]) P[k][c] / (w_i - (E_k - E0[0]) + i gamma/2),
 *   w_i = (i/ne)(emax-emin) + emin — BlockListOfPolesToSpectraWithBroadening
 *   (Core/BasicMathBlock/BasicMath_BlockTypeConversion.cpp:588-660).
 * out: complex. */
qs_status qs_pole_spectrum(qs_ctx *ctx, int npole, int nv, const double *E, const double *P, int tensor,
                           const double *E0, double emin, double emax, int ne, double gamma, double *out);

/* Broadening of nspec spectra of npts points on a grid of step dE, in place
 * (re, im: nspec*npts host arrays): out(o) = sum_i in(i) K(o, i), K the
 * Gaussian (kind 0) or Lorentzian (kind 1) line of full width at half maximum
 * f_i, averaged over the 7 sub-points o - i + j/7, j = -3..3.
 * per_point = 0: f_i = fwhm[0] for all i, and nothing changes when
 * 2*fwhm[0] < dE — BroadenG / BroadenL (Core/Spectral.cpp:958-1044).
 * per_point = 1: f_i = max(fwhm[i], dE/2), npts values — BroadenGvar /
 * BroadenLvar (Core/Spectral.cpp:1046-1131). Evaluated on the GPU; the sum
 * over i runs in a fixed order. */
qs_status qs_broaden(qs_ctx *ctx, int nspec, int npts, double dE, int kind, int per_point, const double *fwhm,
                     double *re, double *im);

/* Dense Hermitian eigen-decomposition on the host (Householder + implicit QL),
 * used for the small projected problems. A: n*n complex column-major;
 * w: n eigenvalues ascending; Z: n*n complex column-major eigenvectors. */
qs_status qs_heev_host(int n, const double *A, double *w, double *Z);
/* Real symmetric tridiagonal T (diagonal d[0..m-1], off-diagonal e[0..m-2]) on
 * the host: the lowest eigenpair (Newton on the Sturm sequence, safeguarded by
 * bisection, + inverse iteration; z has
 * unit norm, largest component positive), and the smallest/largest
 * eigenvalues — the pole range of a Lanczos table
 * (TriDiagonalMatrixToEigenvalues, BasicMath_MatrixTypeConversion.cpp:1538). */
qs_status qs_tridiag_lowest_host(int m, const double *d, const double *e, double *E, double *z);
/* qs_tridiag_lowest_host with a bracket [lower, upper] expected to hold the
 * lowest eigenvalue (e.g. while a Lanczos run is extended: the previous
 * estimate, an upper bound, and that estimate minus a multiple of its
 * residual); each end is checked and ignored when wrong; NaN for none. */
qs_status qs_tridiag_lowest_from(int m, const double *d, const double *e, double lower, double upper, double *E,
                                 double *z);
qs_status qs_tridiag_extremes_host(int m, const double *d, const double *e, double *lo, double *hi);

/* ---- continued-fraction Green's function -------------------------------- */
/* G(w_i) = a0 + b0^2/(z - a1 - b1^2/(z - a2 - ...)), z = w_i + E0 + i Gamma/2,
 * w_i = emin + i*(emax-emin)/ne, i = 0..ne (ne+1 points) — TriDiagonalMatrixToSpectrum
 * (Core/Spectral.cpp:1433-1480). nspec spectra, a: nspec*(n+1), b: nspec*n,
 * E0: nspec (host arrays); re/im: nspec*(ne+1) (host). Evaluated on the GPU. */
qs_status qs_cf_spectrum(qs_ctx *ctx, int nspec, int n, const double *a, const double *b,
                         const double *E0, double emin, double emax, int ne, double gamma,
                         double *re, double *im);

#ifdef __cplusplus
}
#endif
#endif /* QS_H */
