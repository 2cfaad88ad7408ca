"""Thin object layer over libqs handles (bases, operators, bound operators).

Device vectors are torch complex128 tensors on the GPU (PyTorch is used only
for device memory); every computation goes through libqs.
"""
from __future__ import annotations

import ctypes as C
import itertools

import numpy as np
import torch

from . import det as D
from ._lib import (QS_ENOTCONVERGED, QS_HOP_ASSEMBLED, QS_HOP_AUTO, QS_HOP_MATFREE, Context, QSError,
                   default_context, load, ptr)


def _torch_ready(device) -> None:
    """Wait for torch's queued work on `device`'s current stream, where tensors
    handed to the library were written. Only that stream: a device-wide
    synchronize would also wait for other host threads' library streams."""
    torch.cuda.current_stream(device).synchronize()


def _device(ctx: Context):
    return torch.device("cuda", ctx.device)


def restriction_arrays(nf: int, restrictions):
    """[(mask_string, lo, hi), ...] -> (masks uint8 (nres, nbytes), lo int32, hi int32)."""
    nres = len(restrictions)
    nb = D.nbytes(nf)
    masks = np.zeros((max(nres, 1), nb), dtype=np.uint8)
    lo = np.zeros(max(nres, 1), dtype=np.int32)
    hi = np.zeros(max(nres, 1), dtype=np.int32)
    for r, (m, a, b) in enumerate(restrictions):
        masks[r] = D.mask_bytes(m, nf) if isinstance(m, str) else np.asarray(m, dtype=np.uint8)
        lo[r], hi[r] = int(a), int(b)
    return nres, masks, lo, hi


class Basis:
    """Sorted determinant set on the device (qs_basis)."""

    def __init__(self, handle, ctx: Context, restrictions=None):
        self.h = handle
        self.ctx = ctx
        self.lib = ctx.lib
        self.n = int(self.lib.qs_basis_size(handle))
        self.nf = int(self.lib.qs_basis_nf(handle))
        self.nw = int(self.lib.qs_basis_words(handle))
        self.restrictions = restrictions

    @classmethod
    def enumerate(cls, nf: int, restrictions, ctx: Context | None = None) -> "Basis":
        ctx = ctx or default_context()
        nres, masks, lo, hi = restriction_arrays(nf, restrictions)
        h = C.c_void_p()
        ctx.check(ctx.lib.qs_basis_enumerate(ctx.h, nf, nres, ptr(masks), ptr(lo), ptr(hi), C.byref(h)),
                  "qs_basis_enumerate")
        return cls(h, ctx, list(restrictions))

    @staticmethod
    def count(nf: int, restrictions) -> int:
        """Sector size from the rank tables (host only; needs no GPU)."""
        lib = load()
        nres, masks, lo, hi = restriction_arrays(nf, restrictions)
        n = C.c_int64()
        st = lib.qs_basis_count(None, nf, nres, ptr(masks), ptr(lo), ptr(hi), C.byref(n))
        if st != 0:
            raise QSError(f"qs_basis_count failed with status {st}")
        return int(n.value)

    @classmethod
    def from_words(cls, nf: int, words: np.ndarray, ctx: Context | None = None) -> "Basis":
        ctx = ctx or default_context()
        w = np.ascontiguousarray(np.asarray(words, dtype=np.uint64).reshape(-1, D.nwords(nf)))
        h = C.c_void_p()
        ctx.check(ctx.lib.qs_basis_from_words(ctx.h, nf, ptr(w), w.shape[0], C.byref(h)), "qs_basis_from_words")
        return cls(h, ctx)

    @classmethod
    def closure(cls, op: "Op", seed: "Basis", restrictions=None, max_dim: int = 1 << 62) -> "Basis | None":
        """The smallest determinant set holding seed and every target of op's terms
        (inside the restrictions), built on the GPU; None when it passes max_dim."""
        ctx = seed.ctx
        nres, masks, lo, hi = restriction_arrays(seed.nf, restrictions or [])
        h = C.c_void_p()
        st = ctx.lib.qs_basis_closure(ctx.h, op.h, seed.h, nres, ptr(masks), ptr(lo), ptr(hi), int(max_dim),
                                      C.byref(h))
        if st == QS_ENOTCONVERGED:
            return None
        ctx.check(st, "qs_basis_closure")
        return cls(h, ctx, list(restrictions) if restrictions else None)

    @classmethod
    def image(cls, op: "Op", seed: "Basis", restrictions=None, max_dim: int = 1 << 62) -> "Basis | None":
        """Every target of op's terms on seed's determinants (inside the
        restrictions), built on the GPU: the output basis of a push-form O|psi>;
        None when it holds more than max_dim determinants."""
        ctx = seed.ctx
        nres, masks, lo, hi = restriction_arrays(seed.nf, restrictions or [])
        h = C.c_void_p()
        st = ctx.lib.qs_basis_image(ctx.h, op.h, seed.h, nres, ptr(masks), ptr(lo), ptr(hi), int(max_dim),
                                    C.byref(h))
        if st == QS_ENOTCONVERGED:
            return None
        ctx.check(st, "qs_basis_image")
        return cls(h, ctx, list(restrictions) if restrictions else None)

    def words(self) -> np.ndarray:
        out = np.zeros((self.n, self.nw), dtype=np.uint64)
        if self.n:
            self.ctx.check(self.lib.qs_basis_export(self.ctx.h, self.h, ptr(out)), "qs_basis_export")
        return out

    def index(self, words: np.ndarray) -> np.ndarray:
        w = np.ascontiguousarray(np.asarray(words, dtype=np.uint64).reshape(-1, self.nw))
        out = np.zeros(w.shape[0], dtype=np.int64)
        if w.shape[0]:
            self.ctx.check(self.lib.qs_basis_index(self.ctx.h, self.h, ptr(w), w.shape[0], ptr(out)),
                           "qs_basis_index")
        return out

    def zeros(self, k: int = 1) -> torch.Tensor:
        shape = (self.n,) if k == 1 else (k, self.n)
        return torch.zeros(shape, dtype=torch.complex128, device=_device(self.ctx))

    def __del__(self):
        try:
            if self.h:
                self.lib.qs_basis_destroy(self.h)
                self.h = None
        except Exception:
            pass


def product_candidates(nf: int, restrictions):
    """Mode groups that may serve as group A of a product layout: the mode set
    of every restriction that does not cover all modes (mask strings with
    spaces ignored, or byte masks)."""
    out = []
    for m, _lo, _hi in restrictions:
        if isinstance(m, str):
            bits = m.replace(" ", "")
            modes = [i for i, ch in enumerate(bits[:nf]) if ch == "1"]
        else:
            mb = np.asarray(m, dtype=np.uint8)
            modes = [i for i in range(nf) if (int(mb[i >> 3]) >> (7 - (i & 7))) & 1]
        if 0 < len(modes) < nf and modes not in out:
            out.append(modes)
    return out


def pack_terms(terms):
    """[(codes (o[0] first), complex value), ...] -> (len int32, codes uint16, values float64)."""
    terms = list(terms)
    if not terms:
        return np.zeros(0, dtype=np.int32), np.zeros(1, dtype=np.uint16), np.zeros(2, dtype=np.float64)
    cs, vs = zip(*terms)
    lens = np.fromiter(map(len, cs), dtype=np.int32, count=len(cs))
    codes = np.fromiter(itertools.chain.from_iterable(cs), dtype=np.uint16, count=int(lens.sum()))
    if codes.size == 0:
        codes = np.zeros(1, dtype=np.uint16)
    vals = np.ascontiguousarray(np.array(vs, dtype=np.complex128).view(np.float64))
    return lens, codes, vals


class Op:
    """Compiled ladder-term table on the device (qs_op)."""

    def __init__(self, nf: int, terms, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.nf = nf
        lens, codes, vals = pack_terms(terms)
        h = C.c_void_p()
        self.ctx.check(self.lib.qs_op_create(self.ctx.h, nf, len(terms), ptr(lens), ptr(codes), ptr(vals),
                                             C.byref(h)), "qs_op_create")
        self.h = h
        self.nterms = int(self.lib.qs_op_nterms(h))

    def __del__(self):
        try:
            if self.h:
                self.lib.qs_op_destroy(self.h)
                self.h = None
        except Exception:
            pass


class Hop:
    """An operator bound to (row basis, column basis) (qs_hop)."""

    MODES = {"matfree": QS_HOP_MATFREE, "assembled": QS_HOP_ASSEMBLED, "auto": QS_HOP_AUTO}

    def __init__(self, op: Op, rows: Basis, cols: Basis | None = None, mode: str = "auto", budget: int = 0):
        cols = rows if cols is None else cols
        self.op, self.rows, self.cols = op, rows, cols
        self.ctx = op.ctx
        self.lib = self.ctx.lib
        h = C.c_void_p()
        self.ctx.check(self.lib.qs_hop_create(self.ctx.h, op.h, rows.h, cols.h, self.MODES[mode], budget,
                                              C.byref(h)), "qs_hop_create")
        self.h = h

    @classmethod
    def like(cls, op: Op, template: "Hop") -> "Hop":
        """op bound to template's bases, sharing template's assembled stream
        (qs_hop_create_like): op must compile to the same terms as
        template's operator, values aside."""
        self = cls.__new__(cls)
        self.op, self.rows, self.cols = op, template.rows, template.cols
        self.ctx = op.ctx
        self.lib = self.ctx.lib
        h = C.c_void_p()
        self.ctx.check(self.lib.qs_hop_create_like(self.ctx.h, op.h, template.h, C.byref(h)), "qs_hop_create_like")
        self.h = h
        if getattr(template, "amodes", None):
            self.amodes = template.amodes  # the same product layout
        return self

    @classmethod
    def product(cls, op: Op, basis: Basis, amodes=None) -> "Hop":
        """op on the sector `basis` in the product layout (qs_hop_create_product):
        modes `amodes` form group A, the others group B; rows are grouped by
        their B configuration. Without `amodes`, every restriction mask that
        does not cover all modes is tried as group A and the split with the
        fewest estimated vector loads is kept. Vectors stay in canonical order
        at apply/lanczos; `row_order()` gives the layout's own order."""
        if amodes is None:
            best = None
            for amodes_c in product_candidates(basis.nf, basis.restrictions or []):
                try:
                    h = cls.product(op, basis, amodes_c)
                except QSError:
                    continue
                cost = h.product_info()["load_cost"]
                if best is None or cost < best[0]:
                    best = (cost, h)
            if best is None:
                raise QSError("qs_hop_create_product: no restriction group gives a product layout")
            return best[1]
        self = cls.__new__(cls)
        self.op, self.rows, self.cols = op, basis, basis
        self.ctx = op.ctx
        self.lib = self.ctx.lib
        mask = np.zeros(D.nbytes(basis.nf), dtype=np.uint8)
        for i in amodes:
            mask[i >> 3] |= np.uint8(1 << (7 - (i & 7)))
        h = C.c_void_p()
        self.ctx.check(self.lib.qs_hop_create_product(self.ctx.h, op.h, basis.h, ptr(mask), C.byref(h)),
                       "qs_hop_create_product")
        self.h = h
        self.amodes = sorted(int(i) for i in amodes)
        return self

    def product_info(self) -> dict:
        info = np.zeros(8, dtype=np.int64)
        cost = C.c_double()
        self.ctx.check(self.lib.qs_hop_product_info(self.h, ptr(info), C.byref(cost)), "qs_hop_product_info")
        keys = ("groups", "slices", "asets", "entries", "mapped", "reserved", "a_slots", "maps")
        out = dict(zip(keys, (int(v) for v in info)))
        out["load_cost"] = cost.value
        return out

    def row_order(self) -> np.ndarray:
        """Canonical index of every row in the order the binding's kernels keep vectors."""
        out = np.zeros(self.rows.n, dtype=np.int64)
        self.ctx.check(self.lib.qs_hop_row_order(self.ctx.h, self.h, ptr(out)), "qs_hop_row_order")
        return out

    def apply_rows(self, x: torch.Tensor, y: torch.Tensor | None = None) -> torch.Tensor:
        """y = O x with x, y in the binding's own row order (row_order())."""
        assert x.dtype == torch.complex128 and x.is_contiguous() and x.numel() == self.cols.n
        if y is None:
            y = self.rows.zeros()
        _torch_ready(x.device)
        self.ctx.check(self.lib.qs_hop_apply_rows(self.ctx.h, self.h, ptr(x), ptr(y)), "qs_hop_apply_rows")
        return y

    def stats(self) -> dict:
        a, b, c, d, e = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int(), C.c_int64()
        self.ctx.check(self.lib.qs_hop_stats(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d), C.byref(e)),
                       "qs_hop_stats")
        out = {"nnz_eff": a.value, "nnz_stored": b.value, "stream_bytes": c.value, "assembled": bool(d.value),
               "nnz_offdiag": e.value}
        if out["assembled"]:
            nb, bh, es, pk, sm = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int(), C.c_int()
            self.ctx.check(self.lib.qs_hop_layout(self.h, C.byref(nb), C.byref(bh), C.byref(es), C.byref(pk),
                                                  C.byref(sm)), "qs_hop_layout")
            out.update(nblocks=nb.value, block_hits=bh.value, ell_slots=es.value, packed=bool(pk.value),
                       shift_min=sm.value)
        return out

    def diagonal(self) -> torch.Tensor:
        """<d_i|O|d_i> for every row (assembled, rows == cols), a complex128 device vector."""
        d = self.rows.zeros()
        _torch_ready(d.device)
        self.ctx.check(self.lib.qs_hop_diagonal(self.ctx.h, self.h, ptr(d)), "qs_hop_diagonal")
        return d

    def dense(self) -> torch.Tensor:
        """The dense matrix <r|O|c> (rows x cols, complex128, on the device), stored
        column-major: the returned tensor is the transpose view of a contiguous
        (cols, rows) tensor, as qs_dense_poles takes it."""
        M = torch.empty((self.cols.n, self.rows.n), dtype=torch.complex128, device=_device(self.ctx))
        _torch_ready(M.device)
        self.ctx.check(self.lib.qs_hop_dense(self.ctx.h, self.h, ptr(M)), "qs_hop_dense")
        return M.t()

    def apply(self, x: torch.Tensor, y: torch.Tensor | None = None, ctx: Context | None = None) -> torch.Tensor:
        """y = O x. ctx: the context (stream) to run on, when not the one the
        binding was made with (a context serves one host thread at a time)."""
        assert x.dtype == torch.complex128 and x.is_contiguous() and x.numel() == self.cols.n
        if y is None:
            y = self.rows.zeros()
        c = ctx or self.ctx
        _torch_ready(x.device)
        c.check(self.lib.qs_hop_apply(c.h, self.h, ptr(x), ptr(y)), "qs_hop_apply")
        return y

    def lanczos(self, v0: torch.Tensor, nsteps: int):
        a = np.zeros(nsteps + 1)
        b = np.zeros(nsteps)
        _torch_ready(v0.device)
        self.ctx.check(self.lib.qs_lanczos(self.ctx.h, self.h, ptr(v0), nsteps, ptr(a), ptr(b)), "qs_lanczos")
        return a, b

    def lanczos_krylov(self, v0: torch.Tensor, nsteps: int, eps: float = 1.4901161193847656e-10):
        """Lanczos with full re-orthogonalization that keeps its basis. Returns a
        (nsteps+1,), b (nsteps,) and the normalized Krylov vectors (nsteps, n)."""
        a = np.zeros(nsteps + 1)
        b = np.zeros(nsteps)
        V = torch.zeros((nsteps, self.rows.n), dtype=torch.complex128, device=_device(self.ctx))
        _torch_ready(v0.device)
        self.ctx.check(self.lib.qs_lanczos_krylov(self.ctx.h, self.h, ptr(v0), nsteps, float(eps), ptr(a), ptr(b),
                                                  ptr(V)), "qs_lanczos_krylov")
        return a, b, V

    def block_lanczos(self, V0: torch.Tensor, nsteps: int, sv_cut: float = 1.4901161193847656e-08):
        """V0: (k, n) complex128 start vectors. Returns A (nsteps+1, k, k), B (nsteps, k, k), kept (nsteps+1)."""
        V0 = V0.contiguous()
        k = V0.shape[0]
        A = np.zeros((nsteps + 1, k, k), dtype=np.complex128)
        B = np.zeros((nsteps, k, k), dtype=np.complex128)
        kept = np.zeros(nsteps + 1, dtype=np.int32)
        _torch_ready(V0.device)
        self.ctx.check(self.lib.qs_block_lanczos(self.ctx.h, self.h, k, ptr(V0), nsteps, sv_cut, ptr(A), ptr(B),
                                                 ptr(kept)), "qs_block_lanczos")
        return A, B, kept

    def ground_states(self, nstates: int, tol: float = 1e-12, max_krylov: int = 500, max_restarts: int = 100,
                      seed: int = 1, start: torch.Tensor | None = None, krylov_start: int = 100,
                      krylov_step: int = 10, cycle_end: bool = False):
        """Lowest eigenpairs. Returns (E (nstates,), vecs (nstates, n) tensor, variances, iterations).

        The Krylov basis holds nstates + krylov_start vectors before the first
        thick restart and grows by krylov_step + nstates per restart up to
        max_krylov (NKrylovStart / NKrylovStep / NKrylovMax). cycle_end tests
        convergence only when the basis reaches its restart size, as the
        reference's Krylov loop does (Core/BlockLanczos.cpp:5176-5215)."""
        E = np.zeros(nstates)
        var = np.zeros(nstates)
        it = C.c_int()
        vecs = torch.zeros((nstates, self.rows.n), dtype=torch.complex128, device=_device(self.ctx))
        if start is not None:
            start = start.contiguous()
        _torch_ready(vecs.device)
        self.ctx.check(self.lib.qs_ground_states_ex(self.ctx.h, self.h, nstates, tol, krylov_start, krylov_step,
                                                    max_krylov, max_restarts, seed, 1 if cycle_end else 0, ptr(start), ptr(E), ptr(vecs),
                                                    ptr(var), C.byref(it)), "qs_ground_states_ex")
        return E, vecs, var, it.value

    def __del__(self):
        try:
            if self.h:
                self.lib.qs_hop_destroy(self.h)
                self.h = None
        except Exception:
            pass


def apply(op: Op, rows: Basis, cols: Basis, x: torch.Tensor) -> torch.Tensor:
    y = rows.zeros()
    _torch_ready(x.device)
    op.ctx.check(op.lib.qs_apply(op.ctx.h, op.h, rows.h, cols.h, ptr(x), ptr(y)), "qs_apply")
    return y


def expect(op: Op, basis: Basis, x: torch.Tensor) -> complex:
    re, im = C.c_double(), C.c_double()
    _torch_ready(x.device)
    op.ctx.check(op.lib.qs_expect(op.ctx.h, op.h, basis.h, ptr(x), C.byref(re), C.byref(im)), "qs_expect")
    return complex(re.value, im.value)


def dot(a: torch.Tensor, b: torch.Tensor, ctx: Context | None = None) -> complex:
    ctx = ctx or default_context()
    re, im = C.c_double(), C.c_double()
    _torch_ready(a.device)
    ctx.check(ctx.lib.qs_dot(ctx.h, a.numel(), ptr(a), ptr(b), C.byref(re), C.byref(im)), "qs_dot")
    return complex(re.value, im.value)


def cf_spectrum(a: np.ndarray, b: np.ndarray, E0, emin: float, emax: float, ne: int, gamma: float,
                ctx: Context | None = None) -> np.ndarray:
    """Continued fractions on the GPU. a: (nspec, n+1), b: (nspec, n). Returns complex (nspec, ne+1)."""
    ctx = ctx or default_context()
    a = np.ascontiguousarray(np.atleast_2d(a), dtype=np.float64)
    b = np.ascontiguousarray(np.atleast_2d(b), dtype=np.float64)
    nspec, n = b.shape
    e0 = np.ascontiguousarray(np.broadcast_to(np.asarray(E0, dtype=np.float64), (nspec,)))
    re = np.zeros((nspec, ne + 1))
    im = np.zeros((nspec, ne + 1))
    ctx.check(ctx.lib.qs_cf_spectrum(ctx.h, nspec, n, ptr(a), ptr(b), ptr(e0), emin, emax, ne, gamma,
                                     ptr(re), ptr(im)), "qs_cf_spectrum")
    return re + 1j * im


def block_cf_spectrum(A: np.ndarray, B: np.ndarray, kept: np.ndarray, E0: float, emin: float, emax: float, ne: int,
                      gamma: float, ctx: Context | None = None) -> np.ndarray:
    """Tensor continued fraction on the GPU: returns (ne+1, k, k) complex, G[i][r][c] = <X_r|(z-H)^-1|X_c>."""
    ctx = ctx or default_context()
    A = np.ascontiguousarray(A, dtype=np.complex128)
    B = np.ascontiguousarray(B, dtype=np.complex128)
    kept = np.ascontiguousarray(kept, dtype=np.int32)
    n, k = B.shape[0], B.shape[1]
    out = np.zeros((ne + 1, k, k), dtype=np.complex128)
    ctx.check(ctx.lib.qs_block_cf_spectrum(ctx.h, k, n, ptr(A), ptr(B), ptr(kept), E0, emin, emax, ne, gamma,
                                           ptr(out)), "qs_block_cf_spectrum")
    return out


def broaden(G: np.ndarray, dE: float, kind: str, fwhm, ctx: Context | None = None) -> np.ndarray:
    """Gaussian (kind "G") or Lorentzian ("L") broadening on the GPU of spectra
    G (nspec, npts) on a grid of step dE. fwhm: one width for every point, or
    npts widths (one per source point, each at least dE/2). Returns a new array."""
    ctx = ctx or default_context()
    G = np.atleast_2d(np.asarray(G, dtype=np.complex128))
    nspec, npts = G.shape
    f = np.ascontiguousarray(np.atleast_1d(np.asarray(fwhm, dtype=np.float64)))
    if f.size not in (1, npts):
        raise QSError(f"broaden: {f.size} widths for {npts} points")
    re = np.ascontiguousarray(G.real)
    im = np.ascontiguousarray(G.imag)
    ctx.check(ctx.lib.qs_broaden(ctx.h, nspec, npts, float(dE), 0 if kind == "G" else 1, int(f.size == npts and npts > 1),
                                 ptr(f), ptr(re), ptr(im)), "qs_broaden")
    return re + 1j * im


def dense_poles(Hd: torch.Tensor, X: torch.Tensor | None = None, ctx: Context | None = None, vectors: bool = True):
    """Eigen-decomposition of the dense Hermitian matrix Hd (from Hop.dense()) and
    projections of the vectors X (nv, n) on the eigenvectors. Returns E (n,)
    ascending and P (n, nv) with P[k, s] = <z_k|x_s>. vectors: Hd is overwritten
    by the eigenvectors (Hd[:, k] = z_k); otherwise only E and P are made
    (faster; Hd is overwritten by the Householder reduction)."""
    ctx = ctx or default_context()
    n = Hd.shape[0]
    st = Hd.t()
    assert Hd.shape == (n, n) and st.is_contiguous() and Hd.dtype == torch.complex128
    nv = 0 if X is None else X.shape[0]
    if nv:
        X = X.contiguous()
        assert X.shape == (nv, n) and X.dtype == torch.complex128
    E = np.zeros(n)
    P = np.zeros((n, max(nv, 1)), dtype=np.complex128)
    _torch_ready(Hd.device)
    ctx.check(ctx.lib.qs_dense_poles(ctx.h, n, ptr(st), nv, ptr(X) if nv else None, ptr(E), ptr(P), int(vectors)),
              "qs_dense_poles")
    return E, P[:, :nv]


def dense_lowest(Hd: torch.Tensor, k: int, ctx: Context | None = None):
    """All eigenvalues (ascending) of the dense Hermitian matrix Hd (from
    Hop.dense(); overwritten) and the eigenvectors of the k lowest as a (k, n)
    complex128 device tensor."""
    ctx = ctx or default_context()
    n = Hd.shape[0]
    st = Hd.t()
    assert Hd.shape == (n, n) and st.is_contiguous() and Hd.dtype == torch.complex128
    E = np.zeros(n)
    V = torch.empty((max(k, 1), n), dtype=torch.complex128, device=Hd.device)
    _torch_ready(Hd.device)
    ctx.check(ctx.lib.qs_dense_lowest(ctx.h, n, ptr(st), k, ptr(E), ptr(V)), "qs_dense_lowest")
    return E, V[:k]


def pole_spectrum(E: np.ndarray, P: np.ndarray, E0, emin: float, emax: float, ne: int, gamma: float,
                  tensor: bool = False, ctx: Context | None = None) -> np.ndarray:
    """Spectra of pole lists on the GPU. E (npole,), P (npole, nv) = <z_k|x_s>.
    tensor False: (nv, ne+1), G_s(w) = sum_k |P[k,s]|^2 / (w + E0_s - E_k + i gamma/2);
    tensor True: (ne+1, nv, nv), G_rc(w) = sum_k conj(P[k,r]) P[k,c] / (w - (E_k - E0) + i gamma/2)."""
    ctx = ctx or default_context()
    E = np.ascontiguousarray(E, dtype=np.float64)
    P = np.ascontiguousarray(P, dtype=np.complex128)
    npole, nv = P.shape
    e0 = np.ascontiguousarray(np.broadcast_to(np.asarray(E0, dtype=np.float64), (nv,)))
    out = np.zeros((ne + 1, nv, nv) if tensor else (nv, ne + 1), dtype=np.complex128)
    ctx.check(ctx.lib.qs_pole_spectrum(ctx.h, npole, nv, ptr(E), ptr(P), int(tensor), ptr(e0), emin, emax, ne,
                                       gamma, ptr(out)), "qs_pole_spectrum")
    return out


def heev_host(A: np.ndarray):
    """Dense Hermitian eigen-decomposition by libqs' host solver (no GPU needed)."""
    lib = load()
    A = np.asfortranarray(np.asarray(A, dtype=np.complex128))
    n = A.shape[0]
    w = np.zeros(n)
    Z = np.zeros((n, n), dtype=np.complex128, order="F")
    st = lib.qs_heev_host(n, ptr(A), ptr(w), ptr(Z))
    if st != 0:
        raise QSError(f"qs_heev_host failed with status {st}")
    return w, Z


class Batch:
    """Batched Lanczos recurrences advanced in lockstep (qs_batch)."""

    def __init__(self, max_slots: int, ctx: Context | None = None):
        self.cap = max_slots
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        h = C.c_void_p()
        self.ctx.check(self.lib.qs_batch_create(self.ctx.h, max_slots, C.byref(h)), "qs_batch_create")
        self.h = h
        self.depth = {}
        self.n = {}

    def add(self, hop: Hop, v0: torch.Tensor, depth: int, ycoef: np.ndarray | None = None,
            max_depth: int | None = None, keep_basis: bool = False) -> int:
        """keep_basis: the slot keeps its Lanczos vectors ((max_depth+1) * n
        complex on the device) for `ritz`."""
        slot = C.c_int()
        yc = None if ycoef is None else np.ascontiguousarray(ycoef, dtype=np.float64)
        md = depth if max_depth is None else max(depth, max_depth)
        _torch_ready(v0.device)
        flags = 1 if keep_basis else 0
        self.ctx.check(self.lib.qs_batch_add_ex(self.h, hop.h, ptr(v0), depth, md, flags, ptr(yc), C.byref(slot)),
                       "qs_batch_add_ex")
        self.depth[slot.value] = depth
        self.maxdepth = getattr(self, "maxdepth", {})
        self.maxdepth[slot.value] = md
        self.n[slot.value] = hop.rows.n
        return slot.value

    def set_depth(self, slot: int, depth: int):
        self.ctx.check(self.lib.qs_batch_set_depth(self.h, slot, depth), "qs_batch_set_depth")
        self.depth[slot] = depth

    def run(self, steps: int):
        self.ctx.check(self.lib.qs_batch_run(self.h, steps), "qs_batch_run")

    def poll(self, slot: int, want_coef=True, y: torch.Tensor | None = None):
        st = C.c_int()
        d = self.maxdepth[slot]
        a = np.zeros(d + 1) if want_coef else None
        b = np.zeros(d + 1) if want_coef else None
        self.ctx.check(self.lib.qs_batch_poll(self.h, slot, C.byref(st), ptr(a), ptr(b), ptr(y)), "qs_batch_poll")
        return st.value, a, b

    def ritz(self, slot: int, z: np.ndarray, y: torch.Tensor) -> torch.Tensor:
        """y = sum_i z[i] v_i over the slot's kept Lanczos vectors."""
        zz = np.ascontiguousarray(z, dtype=np.float64)
        assert y.dtype == torch.complex128 and y.is_contiguous() and y.numel() == self.n[slot]
        _torch_ready(y.device)
        self.ctx.check(self.lib.qs_batch_ritz(self.h, slot, len(zz), ptr(zz), ptr(y)), "qs_batch_ritz")
        return y

    def remove(self, slot: int):
        self.ctx.check(self.lib.qs_batch_remove(self.h, slot), "qs_batch_remove")
        self.depth.pop(slot, None)
        self.n.pop(slot, None)

    def active(self) -> int:
        return int(self.lib.qs_batch_active(self.h))

    def coeffs(self, slots, ld: int):
        """Completed steps (n,), a (n, ld) and b (n, ld) of several slots in one transfer."""
        sl = np.ascontiguousarray(slots, dtype=np.int32)
        n = len(sl)
        st = np.zeros(n, dtype=np.int32)
        a = np.zeros((n, ld))
        b = np.zeros((n, ld))
        self.ctx.check(self.lib.qs_batch_coeffs(self.h, n, ptr(sl), ld, ptr(st), ptr(a), ptr(b)), "qs_batch_coeffs")
        return st, a, b

    def set_depths(self, slots, depths):
        sl = np.ascontiguousarray(slots, dtype=np.int32)
        dp = np.ascontiguousarray(depths, dtype=np.int32)
        self.ctx.check(self.lib.qs_batch_set_depths(self.h, len(sl), ptr(sl), ptr(dp)), "qs_batch_set_depths")
        for s, d in zip(sl, dp):
            self.depth[int(s)] = int(d)

    def steps(self, cap: int) -> np.ndarray:
        out = np.zeros(cap, dtype=np.int32)
        self.ctx.check(self.lib.qs_batch_steps(self.h, ptr(out)), "qs_batch_steps")
        return out

    def __del__(self):
        try:
            if self.h:
                self.lib.qs_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass


def tridiag_lowest(d: np.ndarray, e: np.ndarray, lower: float | None = None, upper: float | None = None):
    """Lowest eigenpair of the real symmetric tridiagonal matrix (d, e) (host,
    libqs). lower / upper: values expected below / above the lowest eigenvalue,
    which shorten the search (checked; a wrong one is ignored)."""
    lib = load()
    d = np.ascontiguousarray(d, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    m = len(d)
    E = C.c_double()
    z = np.zeros(m)
    if lower is None and upper is None:
        st = lib.qs_tridiag_lowest_host(m, ptr(d), ptr(e) if m > 1 else None, C.byref(E), ptr(z))
    else:
        nan = float("nan")
        st = lib.qs_tridiag_lowest_from(m, ptr(d), ptr(e) if m > 1 else None, nan if lower is None else float(lower),
                                        nan if upper is None else float(upper), C.byref(E), ptr(z))
    if st != 0:
        raise QSError(f"qs_tridiag_lowest_host failed with status {st}")
    return E.value, z


def tridiag_extremes(d: np.ndarray, e: np.ndarray):
    lib = load()
    d = np.ascontiguousarray(d, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    lo, hi = C.c_double(), C.c_double()
    st = lib.qs_tridiag_extremes_host(len(d), ptr(d), ptr(e) if len(d) > 1 else None, C.byref(lo), C.byref(hi))
    if st != 0:
        raise QSError(f"qs_tridiag_extremes_host failed with status {st}")
    return lo.value, hi.value
