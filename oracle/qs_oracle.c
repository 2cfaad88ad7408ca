/*
 * qs_oracle.c — CPU restatement of the reference's second-quantization hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: only tests/, the smoke
 * test and the CPU-baseline leg of bench.py load it. The product (libqs) never
 * links, loads or calls it.
 *
 * It restates, in plain C, the reference algorithm as written (push form,
 * hash-set wavefunctions in insertion order, byte-wise ladder operators):
 *   or_ladder            OperatorCr / OperatorAn / OperatorLader  Core/OperatorWaveFunction.cpp:7-116
 *   or_restricted_ok     QDetNotRestricted                       Core/OperatorWaveFunction.cpp:2260-2286
 *   or_opsi              OperatorPsi{CC} / OperatorPsiRestrictedCC Core/OperatorWaveFunction.cpp:1750, 2288
 *   or_vec_add           ComplexWaveFunctionAddElement            Core/ComplexWaveFunctionArithmetic.cpp:64
 *   or_dot               ComplexWaveFunctionDotProduct            Core/ComplexWaveFunctionArithmetic.cpp:991
 *   or_truncate          ComplexWaveFunctionTruncateBasis         Core/ComplexWaveFunctionArithmetic.cpp:1073
 *   or_lanczos           LanczosTriDiagonalize{,Restricted}CC     Core/Lanczos.cpp:2169-2264, 2501
 *   or_cf                TriDiagonalMatrixToSpectrum              Core/Spectral.cpp:1433-1480
 *   or_broaden           BroadenG / BroadenL / BroadenGvar / BroadenLvar
 *                                                                 Core/Spectral.cpp:958-1131
 *   or_fwhm_table        the {{E, FWHM}, ...} interpolation of LuaSpectraBroaden
 *                                                                 LuaInterface/Spectra.cpp:44-62
 *   or_enumerate         the restricted sector as a set (the reference grows it by
 *                        repeated H*psi, Core/BlockLanczos.cpp:5849-5858); restated
 *                        as a depth-first walk over modes in ascending numeric order.
 * Determinants are the reference's bytes: mode i is bit 128>>(i&7) of byte i>>3.
 * Bosons are not supported (the reference's NewWavefunction rejects NB>0,
 * LuaInterface/ssp.cpp:950-953).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_MAXB 16 /* up to 128 modes */

typedef struct {
  uint8_t b[OR_MAXB];
} ordet;

/* ------------------------------------------------------------------------ */
/* ladder operators                                                          */
/* ------------------------------------------------------------------------ */
static int parity8(uint8_t x) {
  x ^= x >> 4;
  x ^= x >> 2;
  x ^= x >> 1;
  return x & 1;
}

/* c+_i: 0 if occupied, else set the bit and return (-1)^(occupied modes < i) */
static double or_cr(uint8_t *det, int i) {
  const int k = i >> 3, m = i & 7;
  const uint8_t bit = (uint8_t)(128u >> m);
  if (det[k] & bit) return 0.0;
  det[k] = (uint8_t)(det[k] + bit);
  uint8_t acc = (uint8_t)(m ? (det[k] >> (8 - m)) : 0);
  for (int j = 0; j < k; ++j) acc ^= det[j];
  return parity8(acc) ? -1.0 : 1.0;
}

static double or_an(uint8_t *det, int i) {
  const int k = i >> 3, m = i & 7;
  const uint8_t bit = (uint8_t)(128u >> m);
  if (!(det[k] & bit)) return 0.0;
  det[k] = (uint8_t)(det[k] - bit);
  uint8_t acc = (uint8_t)(m ? (det[k] >> (8 - m)) : 0);
  for (int j = 0; j < k; ++j) acc ^= det[j];
  return parity8(acc) ? -1.0 : 1.0;
}

/* o[0] applied first; bit 15 = create */
double or_ladder(uint8_t *det, const uint16_t *o, int len) {
  double phase = 1.0;
  for (int i = 0; i < len; ++i) {
    const int q = o[i] & 0x7FFF;
    phase *= (o[i] & 0x8000) ? or_cr(det, q) : or_an(det, q);
    if (phase == 0.0) break;
  }
  return phase;
}

int or_restricted_ok(const uint8_t *det, int nbytes, int nres, const uint8_t *masks, const int *lo,
                     const int *hi) {
  for (int r = 0; r < nres; ++r) {
    int sum = 0;
    for (int j = 0; j < nbytes; ++j) {
      uint8_t f = (uint8_t)(masks[r * nbytes + j] & det[j]);
      for (int k = 0; k <= 7; ++k) sum += (f >> k) & 1;
    }
    if (sum > hi[r] || sum < lo[r]) return 0;
  }
  return 1;
}

/* ------------------------------------------------------------------------ */
/* wavefunction: hash map determinant -> complex, insertion ordered          */
/* ------------------------------------------------------------------------ */
typedef struct {
  int nbytes;
  int64_t n, cap;
  ordet *det;
  double *re, *im;
  int64_t *slot; /* open addressing: index+1, 0 = empty */
  int64_t nslot;
} orvec;

static uint64_t det_hash(const ordet *d, int nbytes) {
  uint64_t h = 1469598103934665603ULL;
  for (int i = 0; i < nbytes; ++i) {
    h ^= d->b[i];
    h *= 1099511628211ULL;
  }
  h ^= h >> 29;
  h *= 0xbf58476d1ce4e5b9ULL;
  h ^= h >> 32;
  return h;
}

orvec *or_vec_new(int nbytes) {
  orvec *v = (orvec *)calloc(1, sizeof(orvec));
  v->nbytes = nbytes;
  v->cap = 16;
  v->det = (ordet *)calloc((size_t)v->cap, sizeof(ordet));
  v->re = (double *)calloc((size_t)v->cap, sizeof(double));
  v->im = (double *)calloc((size_t)v->cap, sizeof(double));
  v->nslot = 64;
  v->slot = (int64_t *)calloc((size_t)v->nslot, sizeof(int64_t));
  return v;
}

void or_vec_free(orvec *v) {
  if (!v) return;
  free(v->det);
  free(v->re);
  free(v->im);
  free(v->slot);
  free(v);
}

static void vec_rehash(orvec *v) {
  free(v->slot);
  v->nslot *= 2;
  v->slot = (int64_t *)calloc((size_t)v->nslot, sizeof(int64_t));
  for (int64_t i = 0; i < v->n; ++i) {
    uint64_t h = det_hash(&v->det[i], v->nbytes) & (uint64_t)(v->nslot - 1);
    while (v->slot[h]) h = (h + 1) & (uint64_t)(v->nslot - 1);
    v->slot[h] = i + 1;
  }
}

static int64_t vec_find(const orvec *v, const ordet *d) {
  uint64_t h = det_hash(d, v->nbytes) & (uint64_t)(v->nslot - 1);
  while (v->slot[h]) {
    int64_t i = v->slot[h] - 1;
    if (memcmp(v->det[i].b, d->b, (size_t)v->nbytes) == 0) return i;
    h = (h + 1) & (uint64_t)(v->nslot - 1);
  }
  return -1;
}

/* ComplexWaveFunctionAddElement: add to an existing entry or append */
void or_vec_add(orvec *v, const uint8_t *det, double re, double im) {
  ordet d;
  memset(&d, 0, sizeof d);
  memcpy(d.b, det, (size_t)v->nbytes);
  int64_t i = vec_find(v, &d);
  if (i >= 0) {
    v->re[i] += re;
    v->im[i] += im;
    return;
  }
  if (v->n == v->cap) {
    v->cap *= 2;
    v->det = (ordet *)realloc(v->det, (size_t)v->cap * sizeof(ordet));
    v->re = (double *)realloc(v->re, (size_t)v->cap * sizeof(double));
    v->im = (double *)realloc(v->im, (size_t)v->cap * sizeof(double));
  }
  v->det[v->n] = d;
  v->re[v->n] = re;
  v->im[v->n] = im;
  v->n++;
  if (2 * v->n > v->nslot) {
    vec_rehash(v);
  } else {
    uint64_t h = det_hash(&d, v->nbytes) & (uint64_t)(v->nslot - 1);
    while (v->slot[h]) h = (h + 1) & (uint64_t)(v->nslot - 1);
    v->slot[h] = v->n;
  }
}

int64_t or_vec_size(const orvec *v) { return v->n; }

void or_vec_export(const orvec *v, uint8_t *dets, double *vals) {
  for (int64_t i = 0; i < v->n; ++i) {
    memcpy(dets + i * v->nbytes, v->det[i].b, (size_t)v->nbytes);
    vals[2 * i] = v->re[i];
    vals[2 * i + 1] = v->im[i];
  }
}

/* value of one determinant (0 if absent) */
void or_vec_get(const orvec *v, const uint8_t *det, double *re, double *im) {
  ordet d;
  memset(&d, 0, sizeof d);
  memcpy(d.b, det, (size_t)v->nbytes);
  int64_t i = vec_find(v, &d);
  *re = i >= 0 ? v->re[i] : 0.0;
  *im = i >= 0 ? v->im[i] : 0.0;
}

static orvec *vec_copy(const orvec *a) {
  orvec *v = or_vec_new(a->nbytes);
  for (int64_t i = 0; i < a->n; ++i) or_vec_add(v, a->det[i].b, a->re[i], a->im[i]);
  return v;
}

/* <a|b>, a conjugated; lookups into b over a's entries (ComplexWaveFunctionDotProduct) */
void or_dot(const orvec *a, const orvec *b, double *re, double *im) {
  double sr = 0.0, si = 0.0;
  for (int64_t i = 0; i < a->n; ++i) {
    int64_t j = vec_find(b, &a->det[i]);
    if (j < 0) continue;
    sr += a->re[i] * b->re[j] + a->im[i] * b->im[j];
    si += a->re[i] * b->im[j] - a->im[i] * b->re[j];
  }
  *re = sr;
  *im = si;
}

static double vec_norm(const orvec *a) {
  double s = 0.0;
  for (int64_t i = 0; i < a->n; ++i) s += a->re[i] * a->re[i] + a->im[i] * a->im[i];
  return sqrt(s);
}

static void vec_scale(orvec *a, double f) {
  for (int64_t i = 0; i < a->n; ++i) {
    a->re[i] *= f;
    a->im[i] *= f;
  }
}

/* keep entries with |c|^2 > eps^2 (ComplexWaveFunctionTruncateBasis) */
orvec *or_truncate(orvec *a, double eps) {
  const double e2 = eps * eps;
  orvec *v = or_vec_new(a->nbytes);
  for (int64_t i = 0; i < a->n; ++i)
    if (a->re[i] * a->re[i] + a->im[i] * a->im[i] > e2) or_vec_add(v, a->det[i].b, a->re[i], a->im[i]);
  or_vec_free(a);
  return v;
}

/* a += f * b over the union of both bases (ComplexWaveFunctionAddMultipliedUnionBasis) */
static void vec_axpy_union(orvec *a, const orvec *b, double f) {
  for (int64_t i = b->n; i != 0; --i) or_vec_add(a, b->det[i - 1].b, f * b->re[i - 1], f * b->im[i - 1]);
}

/* ------------------------------------------------------------------------ */
/* operators and O*psi (push form)                                           */
/* ------------------------------------------------------------------------ */
typedef struct {
  int nf, nbytes, nterms;
  int *len;
  int64_t *off;
  uint16_t *codes;
  double *re, *im;
} orop;

orop *or_op_new(int nf, int nterms, const int *len, const uint16_t *codes, const double *vals) {
  orop *o = (orop *)calloc(1, sizeof(orop));
  o->nf = nf;
  o->nbytes = (nf + 7) / 8;
  o->nterms = nterms;
  o->len = (int *)malloc(sizeof(int) * (size_t)(nterms + 1));
  o->off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nterms + 1));
  o->re = (double *)malloc(sizeof(double) * (size_t)(nterms + 1));
  o->im = (double *)malloc(sizeof(double) * (size_t)(nterms + 1));
  int64_t tot = 0;
  for (int t = 0; t < nterms; ++t) {
    o->len[t] = len[t];
    o->off[t] = tot;
    tot += len[t];
    o->re[t] = vals[2 * t];
    o->im[t] = vals[2 * t + 1];
  }
  o->codes = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)(tot + 1));
  memcpy(o->codes, codes, sizeof(uint16_t) * (size_t)tot);
  return o;
}

void or_op_free(orop *o) {
  if (!o) return;
  free(o->len);
  free(o->off);
  free(o->codes);
  free(o->re);
  free(o->im);
  free(o);
}

/* Opsi = O psi. With nres > 0 the targets are filtered by the restrictions
 * (OperatorPsiRestrictedCC); without, every target is kept (OperatorPsiCC).
 * Sources are visited in psi's order, terms in the operator's order.
 * nthreads > 1: sources are split in contiguous chunks, each thread fills a
 * private map, and the maps are merged in thread order. */
orvec *or_opsi_range(const orop *O, const orvec *psi, int64_t i0, int64_t i1, int nres,
                     const uint8_t *masks, const int *lo, const int *hi, int nthreads) {
  const int nb = O->nbytes;
  if (nthreads < 1) nthreads = 1;
  orvec **part = (orvec **)calloc((size_t)nthreads, sizeof(orvec *));
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
  {
    int tid = 0, nt = 1;
#ifdef _OPENMP
    tid = omp_get_thread_num();
    nt = omp_get_num_threads();
#endif
    orvec *loc = or_vec_new(nb);
    const int64_t span = i1 - i0;
    const int64_t a = i0 + span * tid / nt, b = i0 + span * (tid + 1) / nt;
    uint8_t det[OR_MAXB];
    for (int64_t i = a; i < b; ++i) {
      const double pr = psi->re[i], pi = psi->im[i];
      for (int t = 0; t < O->nterms; ++t) {
        memcpy(det, psi->det[i].b, (size_t)nb);
        const double ph = or_ladder(det, O->codes + O->off[t], O->len[t]);
        if (ph == 0.0) continue;
        if (nres && !or_restricted_ok(det, nb, nres, masks, lo, hi)) continue;
        const double vr = ph * O->re[t], vi = ph * O->im[t];
        or_vec_add(loc, det, vr * pr - vi * pi, vr * pi + vi * pr);
      }
    }
    part[tid] = loc;
  }
  orvec *out = part[0];
  for (int t = 1; t < nthreads; ++t) {
    if (!part[t]) continue;
    for (int64_t i = 0; i < part[t]->n; ++i) or_vec_add(out, part[t]->det[i].b, part[t]->re[i], part[t]->im[i]);
    or_vec_free(part[t]);
  }
  free(part);
  return out;
}

orvec *or_opsi(const orop *O, const orvec *psi, int nres, const uint8_t *masks, const int *lo, const int *hi) {
  return or_opsi_range(O, psi, 0, psi->n, nres, masks, lo, hi, 1);
}

/* ------------------------------------------------------------------------ */
/* Lanczos tridiagonalization (LanczosTriDiagonalizeCC, Core/Lanczos.cpp:2169-2264) */
/* ------------------------------------------------------------------------ */
/* nthreads: threads of each O*psi (or_opsi_range). keep (N handles, or NULL):
 * receives a copy of every normalized Lanczos vector v_0 .. v_{N-1} (after
 * truncation), for a Ritz vector by or_combine. */
void or_lanczos_mt(const orop *H, const orvec *psi, int N, double epsilon, double gamma, int nres,
                   const uint8_t *masks, const int *lo, const int *hi, double *a, double *b, int nthreads,
                   orvec **keep) {
  orvec *v[3] = {NULL, NULL, NULL};
  double err = 1.0, junk;
  v[0] = vec_copy(psi);
  a[0] = 0.0;
  b[0] = vec_norm(v[0]);
  vec_scale(v[0], 1.0 / b[0]);
  v[0] = or_truncate(v[0], epsilon);
  if (keep) keep[0] = vec_copy(v[0]);
  v[1] = or_opsi_range(H, v[0], 0, v[0]->n, nres, masks, lo, hi, nthreads);
  or_dot(v[1], v[0], &a[1], &junk);
  int i, vi = 0, vip = 0, vim = 0;
  for (i = 1; i < N; ++i) {
    vip = (i + 1) % 3;
    vi = i % 3;
    vim = (i - 1) % 3;
    vec_axpy_union(v[vi], v[vim], -a[i]);
    b[i] = vec_norm(v[vi]);
    vec_scale(v[vi], 1.0 / b[i]);
    v[vi] = or_truncate(v[vi], epsilon / err);
    if (keep) keep[i] = vec_copy(v[vi]);
    err *= b[i] * b[i] / (b[i] * b[i] + 4.0 * gamma * gamma);
    if (v[vip]) or_vec_free(v[vip]);
    v[vip] = or_opsi_range(H, v[vi], 0, v[vi]->n, nres, masks, lo, hi, nthreads);
    or_dot(v[vip], v[vi], &a[i + 1], &junk);
    vec_axpy_union(v[vip], v[vim], -b[i]);
    or_vec_free(v[vim]);
    v[vim] = NULL;
  }
  for (int k = 0; k < 3; ++k)
    if (v[k]) or_vec_free(v[k]);
}

void or_lanczos(const orop *H, const orvec *psi, int N, double epsilon, double gamma, int nres,
                const uint8_t *masks, const int *lo, const int *hi, double *a, double *b) {
  or_lanczos_mt(H, psi, N, epsilon, gamma, nres, masks, lo, hi, a, b, 1, NULL);
}

/* y = sum_i z[i] vs[i] (real coefficients), accumulated in index order into
 * the union of the supports: the Ritz vector of kept Lanczos vectors. */
orvec *or_combine(orvec *const *vs, int n, const double *z) {
  orvec *y = or_vec_new(vs[0]->nbytes);
  for (int i = 0; i < n; ++i) vec_axpy_union(y, vs[i], z[i]);
  return y;
}

/* ------------------------------------------------------------------------ */
/* continued fraction (TriDiagonalMatrixToSpectrum, Core/Spectral.cpp:1433-1480) */
/* ------------------------------------------------------------------------ */
void or_cf(const double *a, const double *b, int N, double E0, double emin, double emax, int ne,
           double gamma, double *re, double *im) {
  const double dE = (emax - emin) / (double)ne;
  for (int i = 0; i <= ne; ++i) {
    const double hnu = ((double)i * dE) + emin + E0;
    double denom = (gamma * 0.5) * (gamma * 0.5) + (hnu - a[N]) * (hnu - a[N]);
    double vr = (hnu - a[N]) / denom;
    double vim = (-gamma * 0.5) / denom;
    for (int j = (N > 1 ? N : 1) - 1; j >= 1; --j) {
      const double b2 = b[j] * b[j];
      denom = (gamma / 2.0 - vim * b2) * (gamma / 2.0 - vim * b2) + (hnu - a[j] - vr * b2) * (hnu - a[j] - vr * b2);
      vr = (hnu - a[j] - vr * b2) / denom;
      vim = (vim * b2 - gamma / 2.0) / denom;
    }
    re[i] = a[0] + vr * b[0] * b[0];
    im[i] = vim * b[0] * b[0];
  }
}

/* ------------------------------------------------------------------------ */
/* restricted sector, ascending numeric order                                */
/* ------------------------------------------------------------------------ */
typedef struct {
  int nf, nbytes, nres;
  const uint8_t *masks;
  const int *lo, *hi;
  int *cnt;
  int *rem; /* rem[r*(nf+1)+i] = mask-r modes at index >= i */
  uint8_t det[OR_MAXB];
  uint8_t *out;
  int64_t n, cap;
} enum_t;

static int in_mask(const enum_t *E, int r, int i) {
  return (E->masks[r * E->nbytes + (i >> 3)] >> (7 - (i & 7))) & 1;
}

static void enum_rec(enum_t *E, int i) {
  for (int r = 0; r < E->nres; ++r) {
    if (E->cnt[r] > E->hi[r]) return;
    if (E->cnt[r] + E->rem[r * (E->nf + 1) + i] < E->lo[r]) return;
  }
  if (i == E->nf) {
    if (E->out && E->n < E->cap) memcpy(E->out + E->n * E->nbytes, E->det, (size_t)E->nbytes);
    E->n++;
    return;
  }
  enum_rec(E, i + 1); /* mode i empty: numerically smaller first */
  E->det[i >> 3] |= (uint8_t)(128u >> (i & 7));
  for (int r = 0; r < E->nres; ++r) E->cnt[r] += in_mask(E, r, i);
  enum_rec(E, i + 1);
  for (int r = 0; r < E->nres; ++r) E->cnt[r] -= in_mask(E, r, i);
  E->det[i >> 3] &= (uint8_t) ~(128u >> (i & 7));
}

/* returns the sector size; writes up to cap determinants (reference bytes) */
int64_t or_enumerate(int nf, int nres, const uint8_t *masks, const int *lo, const int *hi, uint8_t *out,
                     int64_t cap) {
  enum_t E;
  memset(&E, 0, sizeof E);
  E.nf = nf;
  E.nbytes = (nf + 7) / 8;
  E.nres = nres;
  E.masks = masks;
  E.lo = lo;
  E.hi = hi;
  E.out = out;
  E.cap = cap;
  E.cnt = (int *)calloc((size_t)(nres + 1), sizeof(int));
  E.rem = (int *)calloc((size_t)(nres + 1) * (size_t)(nf + 1), sizeof(int));
  for (int r = 0; r < nres; ++r)
    for (int i = nf - 1; i >= 0; --i) E.rem[r * (nf + 1) + i] = E.rem[r * (nf + 1) + i + 1] + in_mask(&E, r, i);
  enum_rec(&E, 0);
  free(E.cnt);
  free(E.rem);
  return E.n;
}

int or_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* broadening (Core/Spectral.cpp:958-1131)                                   */
/* ------------------------------------------------------------------------ */
/* One spectrum of N+1 points (re, im in place), step dE. kind 0 = Gaussian,
 * 1 = Lorentzian. per_point = 0: one FWHM f[0]; nothing happens when
 * 2 f[0] < dE (BroadenG :958, BroadenL :1003). per_point = 1: FWHM f[i] of
 * source point i, raised to at least dE/2 (BroadenGvar :1046, BroadenLvar
 * :1093). Output point Eout sums over source points Ein in ascending order;
 * each weight averages the line shape over Eout - Ein + j/7, j = -3..3. */
void or_broaden(int N, double dE, int kind, int per_point, const double *f, double *re, double *im) {
  if (!per_point && 2 * f[0] < dE) return;
  double *outr = (double *)calloc((size_t)N + 1, sizeof(double));
  double *outi = (double *)calloc((size_t)N + 1, sizeof(double));
#pragma omp parallel for
  for (int eo = 0; eo <= N; ++eo) {
    for (int ei = 0; ei <= N; ++ei) {
      const double fw = per_point ? (f[ei] > dE / 2.0 ? f[ei] : dE / 2.0) : f[0];
      double w = 0;
      if (kind == 0) {
        const double sig = fw / (2.0 * sqrt(2.0 * log(2.0)));
        const double sa = 2.0 * sig * sig;
        const double sb = 1 / (sig * sqrt(2.0 * M_PI) / dE);
        for (int j = -3; j <= 3; ++j) {
          const double x = (eo - ei + j / 7.0) * dE;
          w += exp(-(x * x) / sa);
        }
        w *= sb / 7.0;
      } else {
        for (int j = -3; j <= 3; ++j) {
          const double x = dE * (eo - ei + j / 7.0);
          w += 1 / (fw * fw + 4 * (x * x));
        }
        w *= dE * 2 * fw / (7.0 * M_PI);
      }
      outr[eo] += re[ei] * w;
      outi[eo] += im[ei] * w;
    }
  }
  memcpy(re, outr, ((size_t)N + 1) * sizeof(double));
  memcpy(im, outi, ((size_t)N + 1) * sizeof(double));
  free(outr);
  free(outi);
}

/* FWHM at the grid points emin + j dE, j = 0..N, from a table of nt
 * (energy, FWHM) pairs with ascending energies: linear between the two
 * entries around the point, the first (last) FWHM before (after) the table
 * (LuaInterface/Spectra.cpp:44-62). */
void or_fwhm_table(int nt, const double *E, const double *F, double emin, double dE, int N, double *out) {
  int k = 0;
  for (int j = 0; j <= N; ++j) {
    const double x = emin + j * dE;
    while (k < nt && x > E[k]) ++k;
    if (k == 0)
      out[j] = F[0];
    else if (k == nt)
      out[j] = F[nt - 1];
    else
      out[j] = F[k - 1] + (F[k] - F[k - 1]) * (x - E[k - 1]) / (E[k] - E[k - 1]);
  }
}
