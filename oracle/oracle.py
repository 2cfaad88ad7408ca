"""ctypes front of the CPU restatement (oracle/qs_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke test and the CPU
baseline of bench.py, never by the quanty_amd package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    L = C.CDLL(LIB)
    P, I64, D = C.c_void_p, C.c_int64, C.c_double
    L.or_broaden.argtypes = [C.c_int, D, C.c_int, C.c_int, P, P, P]
    L.or_fwhm_table.argtypes = [C.c_int, P, P, D, D, C.c_int, P]
    L.or_ladder.restype = D
    L.or_ladder.argtypes = [P, P, C.c_int]
    L.or_restricted_ok.restype = C.c_int
    L.or_restricted_ok.argtypes = [P, C.c_int, C.c_int, P, P, P]
    L.or_vec_new.restype = P
    L.or_vec_new.argtypes = [C.c_int]
    L.or_vec_free.argtypes = [P]
    L.or_vec_add.argtypes = [P, P, D, D]
    L.or_vec_size.restype = I64
    L.or_vec_size.argtypes = [P]
    L.or_vec_export.argtypes = [P, P, P]
    L.or_vec_get.argtypes = [P, P, C.POINTER(D), C.POINTER(D)]
    L.or_dot.argtypes = [P, P, C.POINTER(D), C.POINTER(D)]
    L.or_op_new.restype = P
    L.or_op_new.argtypes = [C.c_int, C.c_int, P, P, P]
    L.or_op_free.argtypes = [P]
    L.or_opsi.restype = P
    L.or_opsi.argtypes = [P, P, C.c_int, P, P, P]
    L.or_opsi_range.restype = P
    L.or_opsi_range.argtypes = [P, P, I64, I64, C.c_int, P, P, P, C.c_int]
    L.or_lanczos.argtypes = [P, P, C.c_int, D, D, C.c_int, P, P, P, P, P]
    L.or_lanczos_mt.argtypes = [P, P, C.c_int, D, D, C.c_int, P, P, P, P, P, C.c_int, P]
    L.or_combine.restype = P
    L.or_combine.argtypes = [P, C.c_int, P]
    L.or_cf.argtypes = [P, P, C.c_int, D, D, D, C.c_int, D, P, P]
    L.or_enumerate.restype = I64
    L.or_enumerate.argtypes = [C.c_int, C.c_int, P, P, P, P, I64]
    L.or_max_threads.restype = C.c_int
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data


def _restr(nf, restrictions):
    nb = (nf + 7) // 8
    nres = len(restrictions or [])
    masks = np.zeros((max(nres, 1), nb), dtype=np.uint8)
    lo = np.zeros(max(nres, 1), dtype=np.int32)
    hi = np.zeros(max(nres, 1), dtype=np.int32)
    for r, (m, a, b) in enumerate(restrictions or []):
        bits = [c for c in m if c in "01"]
        assert len(bits) == nf
        for i, c in enumerate(bits):
            if c == "1":
                masks[r, i >> 3] |= 128 >> (i & 7)
        lo[r], hi[r] = a, b
    return nres, masks, lo, hi


class Vec:
    """Reference-style wavefunction (hash map det bytes -> complex, insertion order)."""

    def __init__(self, nf, handle=None):
        self.L = load()
        self.nf = nf
        self.nb = (nf + 7) // 8
        self.h = handle if handle is not None else self.L.or_vec_new(self.nb)

    @classmethod
    def from_arrays(cls, nf, dets_bytes: np.ndarray, vals: np.ndarray) -> "Vec":
        v = cls(nf)
        d = np.ascontiguousarray(dets_bytes, dtype=np.uint8).reshape(-1, v.nb)
        for i in range(d.shape[0]):
            c = complex(vals[i])
            v.L.or_vec_add(v.h, _p(d[i]), c.real, c.imag)
        return v

    def add(self, det_bytes: np.ndarray, val: complex):
        d = np.ascontiguousarray(det_bytes, dtype=np.uint8)
        self.L.or_vec_add(self.h, _p(d), complex(val).real, complex(val).imag)

    def size(self):
        return int(self.L.or_vec_size(self.h))

    def arrays(self):
        n = self.size()
        d = np.zeros((n, self.nb), dtype=np.uint8)
        v = np.zeros(2 * n)
        if n:
            self.L.or_vec_export(self.h, _p(d), _p(v))
        return d, v[0::2] + 1j * v[1::2]

    def dot(self, other: "Vec") -> complex:
        re, im = C.c_double(), C.c_double()
        self.L.or_dot(self.h, other.h, C.byref(re), C.byref(im))
        return complex(re.value, im.value)

    def __del__(self):
        try:
            if self.h:
                self.L.or_vec_free(self.h)
                self.h = None
        except Exception:
            pass


class Op:
    def __init__(self, nf, terms):
        self.L = load()
        self.nf = nf
        lens = np.array([len(c) for c, _ in terms], dtype=np.int32)
        codes = np.array([x for c, _ in terms for x in c] or [0], dtype=np.uint16)
        vals = np.zeros(2 * max(len(terms), 1))
        for i, (_, val) in enumerate(terms):
            vals[2 * i], vals[2 * i + 1] = complex(val).real, complex(val).imag
        self._keep = (lens, codes, vals)
        self.h = self.L.or_op_new(nf, len(terms), _p(lens), _p(codes), _p(vals))

    def apply(self, psi: Vec, restrictions=None, i0=None, i1=None, nthreads=1) -> Vec:
        nres, masks, lo, hi = _restr(self.nf, restrictions)
        i0 = 0 if i0 is None else i0
        i1 = psi.size() if i1 is None else i1
        h = self.L.or_opsi_range(self.h, psi.h, i0, i1, nres, _p(masks), _p(lo), _p(hi), nthreads)
        return Vec(self.nf, h)

    def lanczos(self, psi: Vec, n: int, epsilon=0.0, gamma=0.0, restrictions=None, nthreads=1, keep=False):
        """(a, b) of n Lanczos steps; keep: also the n normalized Lanczos vectors."""
        nres, masks, lo, hi = _restr(self.nf, restrictions)
        a = np.zeros(n + 1)
        b = np.zeros(n)
        hs = (C.c_void_p * n)() if keep else None
        self.L.or_lanczos_mt(self.h, psi.h, n, epsilon, gamma, nres, _p(masks), _p(lo), _p(hi), _p(a), _p(b),
                             int(nthreads), hs)
        if keep:
            return a, b, [Vec(self.nf, C.c_void_p(h)) for h in hs]
        return a, b

    def __del__(self):
        try:
            if self.h:
                self.L.or_op_free(self.h)
                self.h = None
        except Exception:
            pass


def ladder(det_bytes: np.ndarray, codes) -> tuple:
    L = load()
    d = np.array(det_bytes, dtype=np.uint8, copy=True)
    o = np.array(codes or [0], dtype=np.uint16)
    ph = L.or_ladder(_p(d), _p(o), len(codes))
    return ph, d


def enumerate_sector(nf, restrictions) -> np.ndarray:
    L = load()
    nres, masks, lo, hi = _restr(nf, restrictions)
    n = L.or_enumerate(nf, nres, _p(masks), _p(lo), _p(hi), None, 0)
    out = np.zeros((n, (nf + 7) // 8), dtype=np.uint8)
    L.or_enumerate(nf, nres, _p(masks), _p(lo), _p(hi), _p(out), n)
    return out


def cf(a, b, E0, emin, emax, ne, gamma) -> np.ndarray:
    L = load()
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    re = np.zeros(ne + 1)
    im = np.zeros(ne + 1)
    L.or_cf(_p(a), _p(b), len(b), E0, emin, emax, ne, gamma, _p(re), _p(im))
    return re + 1j * im


def broaden(G, dE, kind, fwhm, per_point=False) -> np.ndarray:
    """G: (nspec, npts) complex; kind "G" or "L"; fwhm: a number, or npts values
    when per_point."""
    L = load()
    G = np.atleast_2d(np.asarray(G, dtype=np.complex128))
    f = np.ascontiguousarray(np.atleast_1d(fwhm), dtype=np.float64)
    out = np.empty_like(G)
    for s in range(G.shape[0]):
        re = np.ascontiguousarray(G[s].real)
        im = np.ascontiguousarray(G[s].imag)
        L.or_broaden(G.shape[1] - 1, float(dE), 0 if kind == "G" else 1, int(per_point), _p(f), _p(re), _p(im))
        out[s] = re + 1j * im
    return out


def fwhm_table(table, emin, dE, npts) -> np.ndarray:
    L = load()
    t = np.asarray(table, dtype=np.float64).reshape(-1, 2)
    E = np.ascontiguousarray(t[:, 0])
    F = np.ascontiguousarray(t[:, 1])
    out = np.zeros(npts)
    L.or_fwhm_table(len(t), _p(E), _p(F), float(emin), float(dE), npts - 1, _p(out))
    return out


def max_threads() -> int:
    return int(load().or_max_threads())


def combine(vs, z) -> Vec:
    """sum_i z[i] vs[i] over oracle vectors (real z), in index order."""
    L = load()
    arr = (C.c_void_p * len(vs))(*[v.h for v in vs])
    zz = np.ascontiguousarray(z, dtype=np.float64)
    return Vec(vs[0].nf, C.c_void_p(L.or_combine(arr, len(vs), _p(zz))))
